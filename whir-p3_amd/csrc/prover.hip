// Device-resident WHIR commit/prove behind the C ABI of include/whir_gpu.h.
//
// The host runs the parameter schedule and replays the transcript
// (DuplexChallenger); every pass over a polynomial, codeword or Merkle tree,
// and the transcript steps between them, run on the device and stay resident
// in HBM between steps.
//
// Proofs run in groups: a group is up to GROUP_MAX proofs of one
// configuration (same number of variables and parameters, each with its own
// polynomial, constraints and challenger) that go through the protocol in
// lockstep, every device step one launch for all of them (the proof index is
// the grid's z dimension).  The host side loops over the group's proofs for
// what is per proof (transcript replay, openings, proof objects).  A single
// proof (whirgpu_commit / whirgpu_prove) is a group of one.
#include <chrono>
#include <functional>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/whir_gpu.h"
#include "host.h"
#include "verifier.h"
#include "internal.h"
#include "kernels.h"

using namespace kb;
using whir::Digest;
using whir::Challenger;
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


// One proof's share of a context: reduction scratch, eq tables, the device
// transcript's states and logs, and the host-visible buffers kernels read or
// write directly (small results, query indices, openings, roots, staged
// points).  A context holds one slot per proof of the largest group it has
// run; a single proof uses slot 0.
struct Slot {
  static constexpr int SC_LOG_CAP = 64;
  static constexpr size_t SMALL_CAP = 4096;   // extension elements of h_small
  static constexpr size_t IDX_CAP = 4096;
  void* d_mem = nullptr;   // device memory carved into the buffers below
  void* h_mem = nullptr;   // coherent host memory carved into the host-visible buffers
  size_t d_bytes = 0, h_bytes = 0;
  ef* d_partial = nullptr;   // reduction scratch: partial_cap extension elements
  ef* d_tabs = nullptr;      // eq tables: 8 slots of tab_cap (own allocation, may grow)
  size_t tab_cap = 0;
  // Small reduction results are written by the kernels straight into
  // coherent host memory (h_small; d_small is its device address) and read
  // after a stream synchronization, so no copy is enqueued for them.
  ef* h_small = nullptr, *d_small = nullptr;
  // Query indices in, opened rows out; each holds one call's data and is
  // reused only after the stream has been synchronized.
  u64* h_idx = nullptr, *d_idx = nullptr;
  u32* h_open = nullptr, *d_open = nullptr;
  // Eight root slots, used in turn, so the next round's commitment can be
  // enqueued before this round's root is read.
  u32* h_root = nullptr, *d_root = nullptr;
  int root_next = 0;
  // Device transcript of sumcheck rounds: two challenger states a launch reads
  // one of and writes the other, the grind results (two, 256 bytes apart),
  // the phase's starting challenger (coherent) and the round log (coherent).
  DevChallenger* d_chal = nullptr;
  u32* d_grind = nullptr;
  DevChallenger* h_chal = nullptr, *d_chal_init = nullptr;
  ScLog* h_sclog = nullptr, *d_sclog = nullptr;
  // Device transcript of a round's steps: its log (a WHIR round's, then the
  // final round's), two challenger states, a long grind's base state and the
  // round buffers.
  RoundLog* h_rlog = nullptr, *d_rlog = nullptr;
  DevChallenger* d_rchal = nullptr;
  GrindBase* d_gbase = nullptr;
  RoundBufs rb{};
  // The final polynomial (at most 2^FINAL_MAX_LOG entries) is observed in
  // one staged run and copied to h_small by the step that observes it.
  static constexpr int FINAL_MAX_LOG = 6;
  static_assert((4 << FINAL_MAX_LOG) <= OBS_STAGE, "the final polynomial is observed in one staged run");
  static_assert((1 << FINAL_MAX_LOG) <= SMALL_CAP, "the final polynomial fits the small buffer");
};

struct whirgpu_ctx {
  int device = 0;
  hipStream_t st = nullptr;
  whirgpu_params params{};
  whir::WhirConfig cfg;
  P2C16 c16{};
  P2C24 c24{};
  Tuning tune;           // launch shapes (whirgpu_set_tuning)
  u32* d_tw = nullptr;   // 2 * 4096 twiddles
  Twiddles tw{};
  Pool pool;
  size_t fixed_bytes = 0;   // device bytes of the buffers allocated outside the pool
  size_t tab_cap = 0;       // eq-table slot size for points of the context's size
  size_t partial_cap = 0;
  size_t open_cap = 0;
  std::vector<std::unique_ptr<Slot>> slots;
  // Host-driven grinds: result slots [GRIND_SLOTS][GROUP_MAX] on the device,
  // refilled with p in bulk, and a pinned host copy of one row.
  static constexpr int GRIND_SLOTS = 64;
  u32* d_hgrind = nullptr;
  u32* h_hgrind = nullptr;
  int gslot = GRIND_SLOTS;
  P2C16* d_c16 = nullptr;     // device copies of the constants (kernels that loop over
  P2C24* d_c24 = nullptr;     // permutations read them from memory)
  // Recorded after a round's transcript steps and sumcheck: the host replays
  // them once it has passed while the next commitment runs.
  hipEvent_t ev_round = nullptr;
  int grind_blocks = 256;
  double timings[7] = {0};
  // live kernel timing: HIP events recorded on `st` around each launch group
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
  void* h_pin = nullptr;
  size_t h_pin_bytes = 0;
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
  // for: a part of the ring is reused only after a stream synchronization, by
  // which time every earlier copy has completed.  stage() hands out space in
  // the ring and the matching space of a device mirror, so several uploads
  // go down as one copy (flush()).
  unsigned char* h_ring = nullptr;
  unsigned char* d_ring = nullptr;   // device mirror of h_ring
  size_t ring_cap = 0, ring_pos = 0, ring_flushed = 0;
  // Points of eq tables (32 coordinates each), read by the kernels from
  // coherent host memory; entries are reused only after a synchronization.
  static constexpr int PTS_CAP = 512;
  ef* h_pts = nullptr;
  ef* d_pts = nullptr;   // device address of h_pts
  int pts_next = 0;
  void sync() {
    flush();
    HIPCHK(hipStreamSynchronize(st));
    ring_pos = ring_flushed = 0;
    pts_next = 0;
  }
  void flush() {
    if (ring_pos > ring_flushed)
      HIPCHK(hipMemcpyAsync(d_ring + ring_flushed, h_ring + ring_flushed, ring_pos - ring_flushed,
                            hipMemcpyHostToDevice, st));
    ring_flushed = ring_pos;
  }
  // Space for n elements: the host pointer to fill and the device address it
  // will have after flush().
  template <typename T>
  std::pair<T*, T*> stage(size_t n) {
    const size_t need = (n * sizeof(T) + 255) & ~size_t(255);
    if (need > ring_cap) bad("staged upload larger than the ring", WHIRGPU_BAD_STATE);
    if (ring_pos + need > ring_cap) {
      flush();
      sync();
    }
    T* h = reinterpret_cast<T*>(h_ring + ring_pos);
    T* d = reinterpret_cast<T*>(d_ring + ring_pos);
    ring_pos += need;
    return {h, d};
  }
  // An upload straight from the ring to dst (large ones through pinned
  // memory, waited for).
  template <typename T>
  void h2d(T* dst, const T* src, size_t n) {
    const size_t bytes = n * sizeof(T);
    if (bytes > ring_cap / 4) {
      T* p = pin<T>(n);
      memcpy(p, src, bytes);
      HIPCHK(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, st));
      sync();
      return;
    }
    flush();
    auto sp = stage<T>(n);
    memcpy(sp.first, src, bytes);
    HIPCHK(hipMemcpyAsync(dst, sp.first, bytes, hipMemcpyHostToDevice, st));
    ring_flushed = ring_pos;   // uploaded directly, not through the mirror
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
  bool root_on_host = false;   // the last tree kernel wrote the root to the slot's h_root
  Slot* sl = nullptr;          // the slot of the proof that committed it
  int root_slot = 0;           // at sl->h_root + 8 * root_slot
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
  // Closes the span; returns its record index (SIZE_MAX when not profiling).
  size_t close() {
    if (!open) return SIZE_MAX;
    hipEvent_t b = c->event();
    HIPCHK(hipEventRecord(b, c->st));
    c->recs.push_back({stage, a, b, work});
    open = false;
    return c->recs.size() - 1;
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

int ilog2u(u64 x) {
  int k = 0;
  while ((x >> (k + 1)) != 0) k++;
  return k;
}

// Carves aligned buffers out of one allocation (first pass: base null, sizes only).
struct Carve {
  char* base;
  size_t off = 0;
  template <typename T>
  T* take(size_t n) {
    off = (off + 255) & ~size_t(255);
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += n * sizeof(T);
    return p;
  }
};

// The slot's buffers, in one device and one coherent host allocation (the eq
// tables apart, so they can grow).
void slot_layout(const whirgpu_ctx* c, Slot* sl, char* d, char* h, char* hdev, size_t* d_bytes, size_t* h_bytes) {
  Carve dv{d};
  sl->d_partial = dv.take<ef>(c->partial_cap);
  sl->d_chal = dv.take<DevChallenger>(2);
  sl->d_grind = dv.take<u32>(2 * 64);   // two results, 256 bytes apart
  sl->d_rchal = dv.take<DevChallenger>(2);
  sl->d_gbase = dv.take<GrindBase>(1);
  sl->rb.pts = dv.take<ef>((size_t)RT_MAX_OOD * RT_MAX_VARS);
  sl->rb.ev = dv.take<ef>(RT_MAX_OOD);
  sl->rb.idx = dv.take<u64>(RT_MAX_QUERIES);
  sl->rb.selz = dv.take<u32>(RT_MAX_QUERIES);
  sl->rb.gpow = dv.take<ef>(RT_MAX_OOD + RT_MAX_QUERIES);
  sl->rb.grind = dv.take<u32>(64);
  sl->rb.nq = dv.take<u32>(64);
  *d_bytes = dv.off;
  // host-visible buffers: the same offsets in the host and device views
  Carve hv{h}, dh{hdev};
  sl->h_small = hv.take<ef>(Slot::SMALL_CAP);
  sl->d_small = dh.take<ef>(Slot::SMALL_CAP);
  sl->h_idx = hv.take<u64>(Slot::IDX_CAP);
  sl->d_idx = dh.take<u64>(Slot::IDX_CAP);
  sl->h_open = hv.take<u32>(c->open_cap);
  sl->d_open = dh.take<u32>(c->open_cap);
  sl->h_root = hv.take<u32>(64);
  sl->d_root = dh.take<u32>(64);
  sl->h_chal = hv.take<DevChallenger>(1);
  sl->d_chal_init = dh.take<DevChallenger>(1);
  sl->h_sclog = hv.take<ScLog>(Slot::SC_LOG_CAP);
  sl->d_sclog = dh.take<ScLog>(Slot::SC_LOG_CAP);
  sl->h_rlog = hv.take<RoundLog>(2);
  sl->d_rlog = dh.take<RoundLog>(2);
  *h_bytes = hv.off;
}

void slot_free(Slot* sl) {
  if (sl->d_mem) (void)hipFree(sl->d_mem);
  if (sl->d_tabs) (void)hipFree(sl->d_tabs);
  if (sl->h_mem) (void)hipHostFree(sl->h_mem);
  sl->d_mem = sl->h_mem = nullptr;
  sl->d_tabs = nullptr;
}

void slot_alloc(whirgpu_ctx* c, Slot* sl) {
  size_t db = 0, hb = 0;
  slot_layout(c, sl, nullptr, nullptr, nullptr, &db, &hb);
  HIPCHK(hipMalloc(&sl->d_mem, db));
  sl->d_bytes = db;
  HIPCHK(hipHostMalloc(&sl->h_mem, hb, hipHostMallocCoherent));
  sl->h_bytes = hb;
  void* hdev = nullptr;
  HIPCHK(hipHostGetDevicePointer(&hdev, sl->h_mem, 0));
  slot_layout(c, sl, static_cast<char*>(sl->d_mem), static_cast<char*>(sl->h_mem), static_cast<char*>(hdev), &db,
              &hb);
  sl->tab_cap = c->tab_cap;
  HIPCHK(hipMalloc(&sl->d_tabs, 8 * sl->tab_cap * sizeof(ef)));
  c->fixed_bytes += db + 8 * sl->tab_cap * sizeof(ef);
}

// At least `count` slots (a group of that many proofs).
void ensure_slots(whirgpu_ctx* c, int count) {
  if (count < 1 || count > GROUP_MAX) bad("a group holds 1 to " + std::to_string(GROUP_MAX) + " proofs");
  while ((int)c->slots.size() < count) {
    c->slots.emplace_back(new Slot());
    slot_alloc(c, c->slots.back().get());
  }
}

// Grows a slot's eq-table scratch for points of up to n coordinates.
void ensure_tabs(whirgpu_ctx* c, Slot* sl, int n) {
  const size_t need = (1ull << std::max(0, n - 8)) + 512;
  if (need <= sl->tab_cap) return;
  c->sync();
  ef* t = nullptr;
  HIPCHK(hipMalloc(&t, 8 * need * sizeof(ef)));
  (void)hipFree(sl->d_tabs);
  sl->d_tabs = t;
  c->fixed_bytes += 8 * (need - sl->tab_cap) * sizeof(ef);
  sl->tab_cap = need;
}

// The points of one launch (point b's coordinates zs[b][from..]) in the
// context's coherent point ring, where the kernels read them; returns their
// device addresses.  A launch's points are taken together, so the
// synchronization that recycles a full ring comes before any of them is written.
std::vector<const ef*> stage_points(whirgpu_ctx* c, const std::vector<const std::vector<ef>*>& zs, size_t from = 0) {
  const int count = (int)zs.size();
  if (count > whirgpu_ctx::PTS_CAP) bad("too many points in one launch", WHIRGPU_BAD_STATE);
  if (c->pts_next + count > whirgpu_ctx::PTS_CAP) c->sync();
  std::vector<const ef*> out(count);
  for (int b = 0; b < count; b++) {
    const std::vector<ef>& z = *zs[b];
    if (z.size() < from || z.size() - from > 32) bad("point of more than 32 coordinates");
    ef* h = c->h_pts + (size_t)32 * c->pts_next;
    out[b] = c->d_pts + (size_t)32 * c->pts_next;
    c->pts_next++;
    for (size_t i = from; i < z.size(); i++) h[i - from] = z[i];
  }
  return out;
}

void read_small(whirgpu_ctx* c, const Slot* sl, ef* dst, size_t n) {
  c->sync();
  memcpy(dst, sl->h_small, n * sizeof(ef));
}

// Blocks of a segmented eq-weighted sum: 2048 >> seg_bits, at most one per row.
int dot_blocks(int rows_log, int seg_bits) {
  long rows = 1L << rows_log;
  long want = 2048L >> seg_bits;
  if (want < 1) want = 1;
  return (int)std::min(rows, want);
}

// Segmented eq-weighted sums of each proof's base polynomial at one point
// each: sum_x eq(z_rest, x) f(s, x) for the 2^sb leading-variable segments s,
// into the slot's h_small (read after a synchronization).
void base_partials(whirgpu_ctx* c, const std::vector<const u32*>& polys, const std::vector<Slot*>& sls, int n,
                   int sb, const std::vector<std::vector<ef>>& z_rest) {
  const int B = (int)polys.size();
  const int nrest = n - sb;
  const int lo = std::min(8, nrest), hi = nrest - lo;
  const int nb = dot_blocks(nrest - lo, sb);
  if ((size_t)nb * (1ull << sb) > c->partial_cap) bad("reduction scratch too small");
  if ((size_t)1 << sb > Slot::SMALL_CAP) bad("too many segments");
  std::vector<EqSplitArgs> ea(B);
  std::vector<DotArgs> da(B);
  std::vector<const std::vector<ef>*> zp;
  for (const auto& z : z_rest) zp.push_back(&z);
  const auto dz = stage_points(c, zp);
  for (int b = 0; b < B; b++) {
    Slot* sl = sls[b];
    if ((1ull << hi) + (1ull << lo) > sl->tab_cap) bad("eq table too large for the context");
    ef* tab = sl->d_tabs;
    ea[b] = EqSplitArgs{dz[b], ef_one(), tab, tab + (1ull << hi)};
    EqTables t{};
    t.hi[0] = tab;
    t.lo[0] = tab + (1ull << hi);
    da[b] = DotArgs{polys[b], t, sl->d_partial, sl->d_small};
  }
  KCHK(eq_split(ea.data(), B, hi, lo, c->st));
  Span sp(c, ST_EQDOT, 4.0 * (double)(1ull << n) * B);
  KCHK(eq_dot_base(da.data(), B, n, sb, lo, 1, nb, c->st));
}

// InitialStatement::evaluate (src/whir/constraints/statement/initial.rs:143-152)
// of statement b at point zs[b], for all statements of a context at once (slot
// b for statement b): the segment partial sums are kept for the SVO rounds.
std::vector<ef> statement_evaluate(whirgpu_ctx* c, const std::vector<whirgpu_statement*>& sts,
                                   const std::vector<std::vector<ef>>& zs) {
  const int B = (int)sts.size();
  ensure_slots(c, B);
  const int n = sts[0]->n, sb = sts[0]->seg_bits;
  std::vector<const u32*> polys(B);
  std::vector<Slot*> sls(B);
  std::vector<std::vector<ef>> rest(B);
  for (int b = 0; b < B; b++) {
    if (sts[b]->ctx != c || sts[b]->n != n || sts[b]->seg_bits != sb) bad("statements of one group differ");
    if ((int)zs[b].size() != n) bad("point has the wrong number of coordinates");
    polys[b] = sts[b]->d_poly;
    sls[b] = c->slots[b].get();
    rest[b].assign(zs[b].begin() + sb, zs[b].end());
  }
  base_partials(c, polys, sls, n, sb, rest);
  c->sync();
  std::vector<ef> out(B);
  for (int b = 0; b < B; b++) {
    std::vector<ef> part(sls[b]->h_small, sls[b]->h_small + (1ull << sb));
    ef eval;
    if (sb == 0) {
      eval = part[0];
    } else {
      std::vector<ef> zsb(zs[b].begin(), zs[b].begin() + sb);
      auto eqs = whir::eq_table_host(zsb, ef_one());
      eval = ef_zero();
      for (size_t i = 0; i < part.size(); i++) eval = ef_add(eval, ef_mul(eqs[i], part[i]));
    }
    sts[b]->points.push_back(zs[b]);
    sts[b]->evals.push_back(eval);
    sts[b]->partials.push_back(std::move(part));
    out[b] = eval;
  }
  return out;
}

// Multilinear evaluation of an extension polynomial on one slot (host-driven rounds).
ef ext_mle(whirgpu_ctx* c, Slot* sl, const ef* d_f, int m, const std::vector<ef>& z) {
  const int lo = std::min(8, m), hi = m - lo;
  if ((1ull << hi) + (1ull << lo) > sl->tab_cap) bad("eq table too large for the context");
  ef* tab = sl->d_tabs;
  const EqSplitArgs ea{stage_points(c, {&z})[0], ef_one(), tab, tab + (1ull << hi)};
  KCHK(eq_split(&ea, 1, hi, lo, c->st));
  EqTables t{};
  t.hi[0] = tab;
  t.lo[0] = tab + (1ull << hi);
  const DotArgs da{d_f, t, sl->d_partial, sl->d_small};
  {
    Span sp(c, ST_EQDOT, 16.0 * (double)(1ull << m));
    KCHK(eq_dot_ext(&da, 1, m, 0, lo, 1, dot_blocks(m - lo, 0), c->st));
  }
  ef out;
  read_small(c, sl, &out, 1);
  return out;
}

ef eq_point(const std::vector<ef>& a, const std::vector<ef>& b) { return whir::eq_at(a, b); }

// Blocks per proof of a grind over all of [0, p) (so the search always
// ends): sized to the expected search (16 * 2^bits candidates), one pass of
// it at most, so short grinds leave the device to other work.
int grind_blocks_for(const whirgpu_ctx* c, int bits) {
  const u64 want = ((16ull << bits) + 255) / 256;
  const int all = c->tune.grind_blocks > 0 ? c->tune.grind_blocks : c->grind_blocks;
  return bits <= c->tune.grind_wide_bits ? all : (int)std::max<u64>(1, std::min<u64>((u64)all, want));
}

// Smallest-witness grinding on the device for each challenger (host state),
// one launch for all of them; updates each challenger like
// GrindingChallenger::grind (observe the witness, sample bits) and returns
// the witnesses (Montgomery form).
// GROUP_MAX result words for one grind, set to p (no witness yet); the slots
// are reset together once all have been used.
u32* grind_slot(whirgpu_ctx* c) {
  if (c->gslot == whirgpu_ctx::GRIND_SLOTS) {
    std::vector<u32> init((size_t)whirgpu_ctx::GRIND_SLOTS * GROUP_MAX, P);
    c->h2d(c->d_hgrind, init.data(), init.size());
    c->gslot = 0;
  }
  return c->d_hgrind + (size_t)c->gslot++ * GROUP_MAX;
}

std::vector<u32> gpu_grind(whirgpu_ctx* c, const std::vector<DuplexChallenger*>& chs, int bits) {
  const int B = (int)chs.size();
  if (B < 1 || B > GROUP_MAX) bad("grind group size", WHIRGPU_BAD_STATE);
  u32* res = grind_slot(c);
  std::vector<GrindArgs> a(B);
  for (int b = 0; b < B; b++) {
    const DuplexChallenger& ch = *chs[b];
    memset(&a[b], 0, sizeof a[b]);
    for (int i = 0; i < 16; i++) a[b].base.s[i] = ch.sponge[i];
    for (int i = 0; i < ch.in_len; i++) a[b].base.s[i] = ch.in[i];
    a[b].wpos = ch.in_len;
    a[b].result = res + b;
  }
  Span sp(c, ST_GRIND, 0);
  KCHK(grind_range(c->d_c16, a.data(), B, bits, TR_NONE, 0, P, grind_blocks_for(c, bits), c->tune, c->st));
  const size_t rec = sp.close();
  HIPCHK(hipMemcpyAsync(c->h_hgrind, res, B * sizeof(u32), hipMemcpyDeviceToHost, c->st));
  c->sync();
  std::vector<u32> out(B);
  double work = 0;
  for (int b = 0; b < B; b++) {
    const u32 w = c->h_hgrind[b];
    if (w >= P) bad("no proof-of-work witness below p", WHIRGPU_GRIND_FAILED);
    work += (double)w + 1;
    out[b] = to_monty(w);
    if (!chs[b]->check_witness(bits, out[b])) bad("grinding witness rejected", WHIRGPU_GRIND_FAILED);
  }
  if (rec != SIZE_MAX) c->recs[rec].work = work;
  return out;
}

// Smallest-witness grinding for the Keccak byte challenger on the device
// (k_grind_keccak): the host absorbs the input blocks before the witness's
// bytes, lays out the rest of the input with its padding, and replays the
// found witness on its challenger (which also checks it).
u32 keccak_gpu_grind(whirgpu_ctx* c, Challenger& ch, int bits) {
  constexpr size_t RATE = 136;
  const std::vector<uint8_t>& in = ch.kc.in;
  const size_t L = in.size();
  const size_t M = L + 4;                     // the witness's bytes appended
  const size_t padded = (M / RATE + 1) * RATE;  // at least one padding byte
  const size_t bw = L / RATE;                 // the block holding the witness's first byte
  KeccakGrindArgs a;
  memset(&a, 0, sizeof a);
  auto lane_at = [](const uint8_t* p) {
    uint64_t x = 0;
    for (int k = 0; k < 8; k++) x |= (uint64_t)p[k] << (8 * k);
    return x;
  };
  for (size_t b = 0; b < bw; b++) {
    for (size_t j = 0; j < RATE / 8; j++) a.s0[j] ^= lane_at(&in[b * RATE + 8 * j]);
    keccak_f1600(a.s0);
  }
  a.nblk = (int)(padded / RATE - bw);
  std::vector<uint8_t> tail((size_t)a.nblk * RATE, 0);
  for (size_t i = bw * RATE; i < L; i++) tail[i - bw * RATE] = in[i];
  tail[M - bw * RATE] ^= 0x01;
  tail.back() ^= 0x80;
  for (int j = 0; j < 17; j++) a.s0[j] ^= lane_at(&tail[8 * j]);
  if (a.nblk == 2)
    for (int j = 0; j < 17; j++) a.t1[j] = lane_at(&tail[RATE + 8 * j]);
  const size_t rel = L - bw * RATE;
  a.lo_blk = 0;
  a.lo_lane = (int)(rel / 8);
  a.shift = (int)(8 * (rel % 8));
  const size_t next = (rel / 8 + 1) * 8;   // the lane after, when the bytes run over
  a.hi_blk = a.shift > 32 ? (int)(next / RATE) : a.lo_blk;
  a.hi_lane = a.shift > 32 ? (int)((next % RATE) / 8) : a.lo_lane;
  if (a.hi_blk >= a.nblk) bad("Keccak grind layout", WHIRGPU_BAD_STATE);
  a.result = grind_slot(c);
  const int all = c->tune.grind_blocks > 0 ? c->tune.grind_blocks : c->grind_blocks;
  const int blocks = (int)std::max<u64>(1, std::min<u64>((u64)all, ((4ull << bits) + 255) / 256));
  Span sp(c, ST_GRIND, 0);
  KCHK(keccak_grind(a, bits, blocks, c->st));
  const size_t rec = sp.close();
  HIPCHK(hipMemcpyAsync(c->h_hgrind, a.result, sizeof(u32), hipMemcpyDeviceToHost, c->st));
  c->sync();
  const u32 w = c->h_hgrind[0];
  if (w >= P) bad("no proof-of-work witness below p", WHIRGPU_GRIND_FAILED);
  if (rec != SIZE_MAX) c->recs[rec].work = (double)w + 1;
  const u32 wm = to_monty(w);
  if (!ch.check_witness(bits, wm)) bad("grinding witness rejected", WHIRGPU_GRIND_FAILED);
  return wm;
}

// Grinding for each challenger of a group: on the device for the Poseidon2
// duplex challenger, on the host for the Keccak one.
std::vector<u32> grind_all(whirgpu_ctx* c, const std::vector<Challenger*>& chs, int bits) {
  if (chs[0]->keccak) {
    std::vector<u32> out;
    for (Challenger* ch : chs) out.push_back(keccak_gpu_grind(c, *ch, bits));
    return out;
  }
  std::vector<DuplexChallenger*> dx;
  for (Challenger* ch : chs) dx.push_back(&ch->dx);
  return gpu_grind(c, dx, bits);
}

// Device-resident product polynomial (ProductPolynomial, src/sumcheck/product_polynomial.rs).
struct SumcheckState {
  Buf<ef> e, w;
  int m = 0;
  ef sum = ef_zero();
};

// One proof of a group in flight.
struct Run {
  whirgpu_statement* st = nullptr;
  whir::WhirProof* P = nullptr;
  Challenger ch;
  Slot* sl = nullptr;
  SumcheckState s;
  whirgpu_prover_data* prev = nullptr;   // the commitment the next STIR queries open
  whirgpu_prover_data* cur = nullptr;    // this round's commitment, until it becomes prev
  whirgpu_prover_data* next = nullptr;   // a commitment enqueued ahead of its round
  std::vector<ef> fold_rand;
};

// Round challenge from the round polynomial (h(0), h(2)): the transcript
// (observe, grind, sample) in a proof, or a scripted sequence.
using ChallengeFn = std::function<ef(const ef&, const ef&)>;

// ProductPolynomial::round (product_polynomial.rs:307-364) with host-driven
// challenges, on one slot.
ef sumcheck_round(whirgpu_ctx* c, Slot* sl, SumcheckState& s, const ChallengeFn& next) {
  if (s.m < 1) bad("sumcheck round on a constant polynomial", WHIRGPU_BAD_STATE);
  const u64 h = 1ull << (s.m - 1);
  ScArgs a{};
  a.e = s.e.p;
  a.w = s.w.p;
  a.partial = sl->d_partial;
  a.c0c2 = sl->d_small;
  {
    Span sp(c, ST_SUMCHECK, 64.0 * (double)h);
    KCHK(sc_coeffs(&a, 1, h, c->st));
  }
  ef cc[2];
  read_small(c, sl, cc, 2);
  a.r = next(cc[0], cc[1]);
  {
    Span sp(c, ST_SUMCHECK, 96.0 * (double)h);
    KCHK(sc_fold(&a, 1, h, c->st));
  }
  s.m -= 1;
  s.sum = whir::extrapolate_012(cc[0], ef_sub(s.sum, cc[0]), cc[1], a.r);
  return a.r;
}

// A device-transcript sumcheck phase (launched for a whole group): per round
// the coefficient partials, the grind with its transcript prologue, and the
// fold whose prologue observes the witness and samples r; no host round trip
// until the end.  The host then replays the logged rounds of each proof on
// its own challenger (observe (h(0), h(2)), check the witness, sample) and
// requires every challenge to equal the device's, so the proof and the
// challenger state are exactly those of host-driven rounds
// (SumcheckData::observe_and_sample, src/whir/proof.rs:163-193;
// ProductPolynomial::round, src/sumcheck/product_polynomial.rs:307-364).
struct ScLaunch {
  int rounds = 0;
  int pow_bits = 0;
  int log_base = 0;                            // first ScLog entry of these rounds
  std::vector<const DevChallenger*> out;       // per proof: device challenger after the last round
  std::vector<size_t> grind_recs;              // profile records of the grinds, by round
};

// Uploads a host challenger as the slot's device phase starting state.
void upload_challenger(Slot* sl, const DuplexChallenger& ch) {
  DevChallenger dc{};
  for (int i = 0; i < 16; i++) dc.sponge[i] = ch.sponge[i];
  for (int i = 0; i < ch.in_len; i++) dc.in[i] = ch.in[i];
  for (int i = 0; i < ch.out_len; i++) dc.out[i] = ch.out[i];
  dc.in_len = (u32)ch.in_len;
  dc.out_len = (u32)ch.out_len;
  // h_chal is free: every phase ends with a synchronization before the next
  // one uploads its state.
  *sl->h_chal = dc;
}

// Launches `rounds` device-transcript sumcheck rounds for every proof of the
// group, proof b starting from the challenger at cin[b] (device-visible memory).
ScLaunch sumcheck_launch(whirgpu_ctx* c, const std::vector<Run*>& R, int rounds, int pow_bits,
                         std::vector<const DevChallenger*> cin, int log_base = 0) {
  const int B = (int)R.size();
  ScLaunch L;
  L.rounds = rounds;
  L.pow_bits = pow_bits;
  L.log_base = log_base;
  L.out = cin;
  if (rounds <= 0) return L;
  if (log_base < 0 || log_base + rounds > Slot::SC_LOG_CAP)
    bad("too many sumcheck rounds in one phase", WHIRGPU_BAD_STATE);
  const int m0 = R[0]->s.m;
  if (m0 < rounds) bad("sumcheck round on a constant polynomial", WHIRGPU_BAD_STATE);
  if (pow_bits == 0 && m0 > SC_SMALL_LOG && m0 - rounds < SC_SMALL_LOG) {
    // a launch per round down to 2^SC_SMALL_LOG entries, the rest in one block
    const int k1 = m0 - SC_SMALL_LOG;
    const ScLaunch first = sumcheck_launch(c, R, k1, 0, cin, log_base);
    ScLaunch rest = sumcheck_launch(c, R, rounds - k1, 0, first.out, log_base + k1);
    rest.rounds = rounds;
    rest.log_base = log_base;
    return rest;
  }
  // Each launch reads the challenger from cin and writes the next state to
  // the other buffer; coefficient partials and grind results alternate too,
  // since every block of a launch reads what block 0 of the same launch
  // would otherwise overwrite.
  auto next_state = [&](int b) {
    DevChallenger* ch0 = R[b]->sl->d_chal;
    return cin[b] == ch0 ? ch0 + 1 : ch0;
  };
  auto part = [&](int b, int k) { return R[b]->sl->d_partial + (k ? 4096 : 0); };
  auto gres = [&](int b, int k) { return R[b]->sl->d_grind + (k ? 64 : 0); };
  std::vector<ScArgs> a(B);
  if (pow_bits == 0 && m0 <= SC_SMALL_LOG) {
    // the whole phase in one block per proof
    for (int b = 0; b < B; b++) {
      if (R[b]->s.m != m0) bad("sumcheck sizes of one group differ", WHIRGPU_BAD_STATE);
      a[b] = ScArgs{};
      a[b].e = R[b]->s.e.p;
      a[b].w = R[b]->s.w.p;
      a[b].grind_reset = gres(b, 0);
      ScTr& tr = a[b].tr;
      tr.ch_in = cin[b];
      tr.ch_out = next_state(b);
      tr.log = R[b]->sl->d_sclog + log_base;
      tr.grind_next = gres(b, rounds & 1);
      cin[b] = tr.ch_out;
    }
    {
      Span sp(c, ST_SUMCHECK, 64.0 * (double)(1ull << m0) * B);
      KCHK(sc_small(a.data(), B, m0, rounds, c->d_c16, c->st));
    }
    for (Run* r : R) r->s.m -= rounds;
    L.out = cin;
    return L;
  }
  if (pow_bits > 0 && m0 <= SC_SMALL_LOG) {
    // Per round one block per proof (the previous round's transcript step and
    // fold, this round's coefficients and grinding state), then the search
    // from that state; the last round's step and fold close the phase.
    std::vector<GrindArgs> ga(B);
    for (int i = 0; i < rounds; i++) {
      for (int b = 0; b < B; b++) {
        if (R[b]->s.m != m0) bad("sumcheck sizes of one group differ", WHIRGPU_BAD_STATE);
        DevChallenger* cout = next_state(b);
        a[b] = ScArgs{};
        a[b].e = R[b]->s.e.p;
        a[b].w = R[b]->s.w.p;
        ScTr& tr = a[b].tr;
        tr.ch_in = cin[b];
        tr.ch_out = cout;
        tr.log = R[b]->sl->d_sclog + log_base + i;
        tr.grind = i > 0 ? gres(b, (i - 1) & 1) : nullptr;
        tr.grind_next = gres(b, i & 1);
        tr.pow_bits = pow_bits;
        tr.gbase = R[b]->sl->d_gbase;
        cin[b] = cout;
        memset(&ga[b], 0, sizeof ga[b]);
        ga[b].result = gres(b, i & 1);
        ga[b].tr.gbase = R[b]->sl->d_gbase;
        ga[b].tr.pow_bits = pow_bits;
      }
      {
        Span sp(c, ST_SUMCHECK, 64.0 * (double)(1ull << (m0 - i)) * B);
        KCHK(sc_step(a.data(), B, i > 0 ? m0 - i + 1 : m0, i > 0 ? 1 : 0, c->d_c16, c->st));
      }
      Span sp(c, ST_GRIND, 0);
      KCHK(grind_range(c->d_c16, ga.data(), B, pow_bits, TR_GBASE, 0, P, grind_blocks_for(c, pow_bits), c->tune,
                       c->st));
      L.grind_recs.push_back(sp.close());
    }
    for (int b = 0; b < B; b++) {
      DevChallenger* cout = next_state(b);
      a[b] = ScArgs{};
      a[b].e = R[b]->s.e.p;
      a[b].w = R[b]->s.w.p;
      a[b].partial = part(b, 1);
      ScTr& tr = a[b].tr;
      tr.ch_in = cin[b];
      tr.ch_out = cout;
      tr.log = R[b]->sl->d_sclog + log_base + rounds - 1;
      tr.partial = part(b, 0);
      tr.grind = gres(b, (rounds - 1) & 1);
      tr.grind_next = gres(b, rounds & 1);
      tr.pow_bits = pow_bits;
      cin[b] = cout;
    }
    {
      const u64 h = 1ull << (m0 - rounds);
      Span sp(c, ST_SUMCHECK, 96.0 * (double)h * B);
      unsigned nb_last = 0;
      KCHK(sc_fold_tr(a.data(), B, h, TR_WITNESS_SAMPLE, c->d_c16, &nb_last, c->st));
    }
    for (Run* r : R) r->s.m -= rounds;
    L.out = cin;
    return L;
  }
  int pp = 0;
  unsigned nb = 0;
  {
    const u64 h = 1ull << (m0 - 1);
    for (int b = 0; b < B; b++) {
      if (R[b]->s.m != m0) bad("sumcheck sizes of one group differ", WHIRGPU_BAD_STATE);
      a[b] = ScArgs{};
      a[b].e = R[b]->s.e.p;
      a[b].w = R[b]->s.w.p;
      a[b].partial = part(b, 0);
      a[b].grind_reset = gres(b, 0);
    }
    Span sp(c, ST_SUMCHECK, 64.0 * (double)h * B);
    KCHK(sc_partials(a.data(), B, h, &nb, c->st));
  }
  std::vector<GrindArgs> ga(B);
  for (int i = 0; i < rounds; i++) {
    const u64 h = 1ull << (m0 - i - 1);
    int fold_mode = TR_COEFFS_SAMPLE;
    if (pow_bits > 0) {
      for (int b = 0; b < B; b++) {
        DevChallenger* cout = next_state(b);
        memset(&ga[b], 0, sizeof ga[b]);
        ScTr& tr = ga[b].tr;
        tr.ch_in = cin[b];
        tr.ch_out = cout;
        tr.log = R[b]->sl->d_sclog + log_base + i;
        tr.partial = part(b, pp);
        tr.grind = gres(b, i & 1);
        tr.nb = (int)nb;
        tr.pow_bits = pow_bits;
        tr.gbase = R[b]->sl->d_gbase;
        ga[b].result = tr.grind;
        cin[b] = cout;
      }
      Span sp(c, ST_GRIND, 0);
      KCHK(grind_range(c->d_c16, ga.data(), B, pow_bits, TR_COEFFS_GRIND, 0, P, grind_blocks_for(c, pow_bits),
                       c->tune, c->st));
      L.grind_recs.push_back(sp.close());
      fold_mode = TR_WITNESS_SAMPLE;
    }
    for (int b = 0; b < B; b++) {
      DevChallenger* cout = next_state(b);
      a[b] = ScArgs{};
      a[b].e = R[b]->s.e.p;
      a[b].w = R[b]->s.w.p;
      a[b].partial = part(b, pp ^ 1);
      ScTr& tr = a[b].tr;
      tr.ch_in = cin[b];
      tr.ch_out = cout;
      tr.log = R[b]->sl->d_sclog + log_base + i;
      tr.partial = part(b, pp);
      tr.grind = gres(b, i & 1);
      tr.grind_next = gres(b, (i + 1) & 1);
      tr.nb = (int)nb;
      tr.pow_bits = pow_bits;
      cin[b] = cout;
    }
    {
      Span sp(c, ST_SUMCHECK, 96.0 * (double)h * B);
      KCHK(sc_fold_tr(a.data(), B, h, fold_mode, c->d_c16, &nb, c->st));
    }
    pp ^= 1;
  }
  for (Run* r : R) r->s.m -= rounds;
  L.out = cin;
  return L;
}

// After the phase has completed: replays proof b's logged rounds on its host
// challenger and returns the challenges.
std::vector<ef> sumcheck_replay(whirgpu_ctx* c, Run& r, const ScLaunch& la, whir::SumcheckData& sd) {
  std::vector<ef> rs;
  const int rounds = la.rounds, pow_bits = la.pow_bits;
  std::vector<ScLog> log(r.sl->h_sclog + la.log_base, r.sl->h_sclog + la.log_base + rounds);
  for (int i = 0; i < rounds; i++) {
    const ScLog& L = log[i];
    if (L.err) bad("device transcript: no valid proof-of-work witness", WHIRGPU_GRIND_FAILED);
    sd.polynomial_evaluations.push_back({L.c0, L.c2});
    r.ch.observe_ef(L.c0);
    r.ch.observe_ef(L.c2);
    if (pow_bits > 0) {
      if (!r.ch.check_witness(pow_bits, L.witness)) bad("device transcript: witness rejected", WHIRGPU_GRIND_FAILED);
      sd.pow_witnesses.push_back(L.witness);
      if (i < (int)la.grind_recs.size() && la.grind_recs[i] != SIZE_MAX)
        c->recs[la.grind_recs[i]].work += (double)from_monty(L.witness) + 1;
    }
    const ef x = r.ch.sample_ef();
    if (!ef_eq(x, L.r)) bad("device transcript diverged from the host challenger", WHIRGPU_BAD_STATE);
    r.s.sum = whir::extrapolate_012(L.c0, ef_sub(r.s.sum, L.c0), L.c2, x);
    rs.push_back(x);
  }
  return rs;
}

// `rounds` device-transcript sumcheck rounds of each proof, from its host
// challenger's state; replayed into each proof's SumcheckData `sd(run)`.
void device_sumcheck(whirgpu_ctx* c, const std::vector<Run*>& R, int rounds, int pow_bits,
                     const std::function<whir::SumcheckData&(Run&)>& sd) {
  if (rounds <= 0) return;
  std::vector<const DevChallenger*> cin;
  for (Run* r : R) {
    upload_challenger(r->sl, r->ch.dx);
    cin.push_back(r->sl->d_chal_init);
  }
  const ScLaunch la = sumcheck_launch(c, R, rounds, pow_bits, cin);
  c->sync();
  for (Run* r : R) r->fold_rand = sumcheck_replay(c, *r, la, sd(*r));
}

// Adds sum_i scale_i * eq(z_i, x) to each proof's weights (the same number of
// points for every proof), four points per launch.
void add_eq_weights(whirgpu_ctx* c, const std::vector<Run*>& R, const std::vector<std::vector<std::vector<ef>>>& pts,
                    const std::vector<std::vector<ef>>& scales, bool accumulate_first) {
  const int B = (int)R.size();
  const int m = R[0]->s.m;
  const int lo = std::min(8, m), hi = m - lo;
  const size_t K = pts[0].size();
  std::vector<CombineArgs> ca(B);
  if (K == 0 && !accumulate_first) {
    for (int b = 0; b < B; b++) ca[b] = CombineArgs{R[b]->s.w.p, EqTables{}, nullptr};
    Span sp(c, ST_COMBINE, 16.0 * (double)(1ull << m) * B);
    KCHK(eq_combine(ca.data(), B, 1ull << m, lo, 0, 0, c->st));
    return;
  }
  std::vector<EqSplitArgs> ea(B);
  for (size_t b0 = 0; b0 < K; b0 += 4) {
    const int kk = (int)std::min<size_t>(4, K - b0);
    for (int b = 0; b < B; b++) {
      if (pts[b].size() != K) bad("constraint counts of one group differ");
      ca[b] = CombineArgs{R[b]->s.w.p, EqTables{}, nullptr};
    }
    for (int k = 0; k < kk; k++) {
      std::vector<const std::vector<ef>*> zp;
      for (int b = 0; b < B; b++) zp.push_back(&pts[b][b0 + k]);
      const auto dz = stage_points(c, zp);
      for (int b = 0; b < B; b++) {
        Slot* sl = R[b]->sl;
        if ((1ull << hi) + (1ull << lo) > sl->tab_cap) bad("eq table too large for the context");
        ef* base = sl->d_tabs + (size_t)k * sl->tab_cap;
        ea[b] = EqSplitArgs{dz[b], scales[b][b0 + k], base, base + (1ull << hi)};
        ca[b].tabs.hi[k] = base;
        ca[b].tabs.lo[k] = base + (1ull << hi);
      }
      KCHK(eq_split(ea.data(), B, hi, lo, c->st));
    }
    const bool acc = b0 > 0 || accumulate_first;
    Span sp(c, ST_COMBINE, (acc ? 32.0 : 16.0) * (double)(1ull << m) * B);
    KCHK(eq_combine(ca.data(), B, 1ull << m, lo, kk, acc ? 1 : 0, c->st));
  }
}

// Round i of the initial sumcheck by the split route (see initial_sumcheck):
// (h_i(0), h_i(2)) from the partial sums P (already folded by r_<i).
void svo_round_poly(const whirgpu_statement* st, const std::vector<std::vector<ef>>& P, const std::vector<ef>& ap,
                    const std::vector<ef>& rs, int i, int l0, ef* c0, ef* c2) {
  *c0 = ef_zero();
  *c2 = ef_zero();
  for (size_t j = 0; j < st->points.size(); j++) {
    const auto& z = st->points[j];
    const size_t half = P[j].size() / 2;
    std::vector<ef> z_after(z.begin() + i + 1, z.begin() + l0);
    auto eqa = whir::eq_table_host(z_after, ef_one());
    ef A0 = ef_zero(), A1 = ef_zero();
    for (size_t q = 0; q < half; q++) {
      A0 = ef_add(A0, ef_mul(eqa[q], P[j][q]));
      A1 = ef_add(A1, ef_mul(eqa[q], P[j][half + q]));
    }
    const ef A2 = ef_sub(ef_dbl(A1), A0);
    std::vector<ef> z_before(z.begin(), z.begin() + i);
    const ef pre = eq_point(z_before, rs);
    const ef& zi = z[i];
    const ef e0 = ef_sub(ef_one(), zi);
    const ef e2 = ef_sub(ef_add(ef_dbl(zi), zi), ef_one());
    *c0 = ef_add(*c0, ef_mul(ap[j], ef_mul(pre, ef_mul(e0, A0))));
    *c2 = ef_add(*c2, ef_mul(ap[j], ef_mul(pre, ef_mul(e2, A2))));
  }
}

void svo_fold_partials(std::vector<std::vector<ef>>& P, const ef& r) {
  for (auto& Pj : P) {
    const size_t half = Pj.size() / 2;
    for (size_t q = 0; q < half; q++) Pj[q] = ef_add(Pj[q], ef_mul(r, ef_sub(Pj[half + q], Pj[q])));
    Pj.resize(half);
  }
}

// After the SVO rounds: f folded by all l0 challenges at once, and the weights
// W(x) = sum_j alpha^j eq(z_j[..l0], rs) eq(z_j[l0..], x), for every proof.
void svo_finish(whirgpu_ctx* c, const std::vector<Run*>& R, int l0, const std::vector<std::vector<ef>>& aps,
                const std::vector<std::vector<ef>>& rss) {
  const int B = (int)R.size();
  const int n = R[0]->st->n;
  std::vector<CompressArgs> ca(B);
  // every proof's eq table of rs in one staged upload (one ring allocation,
  // so a wrap of the ring cannot separate them)
  const size_t ns = 1ull << l0;
  auto sp = c->stage<ef>(ns * B);
  for (int b = 0; b < B; b++) {
    Run& r = *R[b];
    r.s.m = n - l0;
    r.s.e = c->alloc<ef>(1ull << r.s.m);
    r.s.w = c->alloc<ef>(1ull << r.s.m);
    const auto eqrs = whir::eq_table_host(rss[b], ef_one());
    memcpy(sp.first + ns * b, eqrs.data(), ns * sizeof(ef));
    ca[b] = CompressArgs{r.st->d_poly, sp.second + ns * b, r.s.e.p};
  }
  c->flush();
  {
    Span sp(c, ST_COMPRESS, (4.0 * (double)(1ull << n) + 16.0 * (double)(1ull << (n - l0))) * B);
    KCHK(compress_multi(ca.data(), B, n, l0, c->st));
  }
  std::vector<std::vector<std::vector<ef>>> rests(B);
  std::vector<std::vector<ef>> scales(B);
  for (int b = 0; b < B; b++) {
    const whirgpu_statement* st = R[b]->st;
    for (size_t j = 0; j < st->points.size(); j++) {
      const auto& z = st->points[j];
      std::vector<ef> zs(z.begin(), z.begin() + l0);
      rests[b].emplace_back(z.begin() + l0, z.end());
      scales[b].push_back(ef_mul(aps[b][j], eq_point(zs, rss[b])));
    }
  }
  add_eq_weights(c, R, rests, scales, false);
}

// Initial sumcheck from the base evaluations (Sumcheck::from_base_evals,
// src/sumcheck/sumcheck_prover.rs:237-292) for every proof of the group.
// Classic and SVO produce the same transcript (asserted by the reference at
// src/sumcheck/tests.rs:473-487); the device always takes the split route
// when n > l0: per constraint, the partial sums P(s) = sum_x eq(z_rest, x)
// f(s, x) over the l0 leading variables s (computed when the constraint was
// evaluated) give every round polynomial
//   h_i(X) = eq(z_<i, r_<i) eq(z_i, X) sum_s' eq(z_>i, s') P~(r_<i, X, s'),
// and afterwards f is folded by all l0 challenges at once.  The rounds'
// proofs of work are one grind launch for the group.
void initial_sumcheck(whirgpu_ctx* c, const std::vector<Run*>& R, int l0, int pow_bits) {
  const int B = (int)R.size();
  std::vector<std::vector<ef>> aps(B), rss(B);
  std::vector<std::vector<std::vector<ef>>> Ps(B);
  for (int b = 0; b < B; b++) {
    Run& r = *R[b];
    const ef alpha = r.ch.sample_ef();
    aps[b] = whir::powers(alpha, r.st->points.size());
    r.s.sum = ef_zero();
    for (size_t j = 0; j < r.st->points.size(); j++) r.s.sum = ef_add(r.s.sum, ef_mul(aps[b][j], r.st->evals[j]));
    Ps[b] = r.st->partials;
  }
  const whirgpu_statement* st0 = R[0]->st;
  if (l0 > 0 && st0->seg_bits == l0 && st0->n > l0 && !R[0]->ch.keccak && l0 <= SVO_DEV_MAX &&
      (int)st0->points.size() <= SVO_MAX_J && (int)st0->points.size() * l0 <= SVO_MAX_Z) {
    // the rounds on the device transcript: per round one block per proof
    // (the previous round's witness step, this round's polynomial and
    // grinding state), then the search; one synchronization, then the host
    // replays the log on its challengers
    const int J = (int)st0->points.size();
    const size_t per = ((size_t)J << l0) + (size_t)J * l0 + (size_t)J;
    auto sp = c->stage<ef>(per * B);
    std::vector<const DevChallenger*> cin(B);
    std::vector<SvoArgs> sa(B);
    for (int b = 0; b < B; b++) {
      Run& r = *R[b];
      if ((int)r.st->points.size() != J || r.st->seg_bits != l0) bad("constraint counts of one group differ");
      ef* h = sp.first + per * b;
      for (int j = 0; j < J; j++) {
        if (r.st->partials[j].size() != (1ull << l0)) bad("internal: split partial sums", WHIRGPU_BAD_STATE);
        memcpy(h + ((size_t)j << l0), r.st->partials[j].data(), sizeof(ef) << l0);
        memcpy(h + ((size_t)J << l0) + (size_t)j * l0, r.st->points[j].data(), sizeof(ef) * l0);
      }
      memcpy(h + ((size_t)J << l0) + (size_t)J * l0, aps[b].data(), sizeof(ef) * J);
      const ef* d = sp.second + per * b;
      sa[b] = SvoArgs{d, d + ((size_t)J << l0), d + ((size_t)J << l0) + (size_t)J * l0, r.sl->d_partial, ScTr{}};
      upload_challenger(r.sl, r.ch.dx);
      cin[b] = r.sl->d_chal_init;
    }
    c->flush();
    ScLaunch la;
    la.rounds = l0;
    la.pow_bits = pow_bits;
    std::vector<GrindArgs> ga(B);
    for (int i = 0; i <= l0; i++) {
      if (i == l0 && pow_bits == 0) break;
      for (int b = 0; b < B; b++) {
        Slot* sl = R[b]->sl;
        DevChallenger* cout = cin[b] == sl->d_chal ? sl->d_chal + 1 : sl->d_chal;
        ScTr& tr = sa[b].tr;
        tr = ScTr{};
        tr.ch_in = cin[b];
        tr.ch_out = cout;
        tr.log = sl->d_sclog;
        tr.grind = i > 0 ? sl->d_grind + ((i - 1) & 1) * 64 : nullptr;
        tr.grind_next = i < l0 ? sl->d_grind + (i & 1) * 64 : nullptr;
        tr.pow_bits = pow_bits;
        tr.gbase = sl->d_gbase;
        cin[b] = cout;
        memset(&ga[b], 0, sizeof ga[b]);
        ga[b].result = sl->d_grind + (i & 1) * 64;
        ga[b].tr.gbase = sl->d_gbase;
        ga[b].tr.pow_bits = pow_bits;
      }
      KCHK(svo_step(sa.data(), B, J, l0, i, pow_bits, c->d_c16, c->st));
      if (i < l0 && pow_bits > 0) {
        Span gs(c, ST_GRIND, 0);
        KCHK(grind_range(c->d_c16, ga.data(), B, pow_bits, TR_GBASE, 0, P, grind_blocks_for(c, pow_bits), c->tune,
                         c->st));
        la.grind_recs.push_back(gs.close());
      }
    }
    c->sync();
    // the fold and weights are enqueued with the logged challenges first, so
    // the device works while the host replays (which requires every logged
    // challenge to equal its own and fails the proof otherwise)
    for (int b = 0; b < B; b++)
      for (int i = 0; i < l0; i++) rss[b].push_back(R[b]->sl->h_sclog[i].r);
    svo_finish(c, R, l0, aps, rss);
    for (int b = 0; b < B; b++) {
      const std::vector<ef> rs = sumcheck_replay(c, *R[b], la, R[b]->P->initial_sumcheck);
      for (int i = 0; i < l0; i++)
        if (!ef_eq(rs[i], rss[b][i])) bad("device transcript diverged from the host challenger", WHIRGPU_BAD_STATE);
      R[b]->fold_rand = rss[b];
    }
    return;
  }
  if (l0 > 0 && st0->seg_bits == l0 && st0->n > l0) {
    std::vector<Challenger*> chs;
    for (Run* r : R) chs.push_back(&r->ch);
    for (int i = 0; i < l0; i++) {
      for (int b = 0; b < B; b++) {
        Run& r = *R[b];
        ef c0, c2;
        svo_round_poly(r.st, Ps[b], aps[b], rss[b], i, l0, &c0, &c2);
        r.P->initial_sumcheck.polynomial_evaluations.push_back({c0, c2});
        r.ch.observe_ef(c0);
        r.ch.observe_ef(c2);
      }
      if (pow_bits > 0) {
        const auto ws = grind_all(c, chs, pow_bits);
        for (int b = 0; b < B; b++) R[b]->P->initial_sumcheck.pow_witnesses.push_back(ws[b]);
      }
      for (int b = 0; b < B; b++) {
        Run& r = *R[b];
        const auto& pe = r.P->initial_sumcheck.polynomial_evaluations.back();
        const ef x = r.ch.sample_ef();
        r.s.sum = whir::extrapolate_012(pe[0], ef_sub(r.s.sum, pe[0]), pe[1], x);
        rss[b].push_back(x);
        svo_fold_partials(Ps[b], x);
      }
    }
    svo_finish(c, R, l0, aps, rss);
    for (int b = 0; b < B; b++) R[b]->fold_rand = rss[b];
    return;
  }
  // small instances (n <= l0): classic rounds on the extension-embedded
  // evaluations, one proof at a time
  for (int b = 0; b < B; b++) {
    Run& r = *R[b];
    const int n = r.st->n;
    r.s.m = n;
    r.s.e = c->alloc<ef>(1ull << n);
    r.s.w = c->alloc<ef>(1ull << n);
    const ExtArgs xa{r.st->d_poly, r.s.e.p};
    KCHK(base_to_ext(&xa, 1, 1ull << n, c->st));
    add_eq_weights(c, {&r}, {r.st->points}, {aps[b]}, false);
    r.fold_rand.clear();
    for (int i = 0; i < l0; i++) {
      r.fold_rand.push_back(sumcheck_round(c, r.sl, r.s, [&](const ef& c0, const ef& c2) {
        r.P->initial_sumcheck.polynomial_evaluations.push_back({c0, c2});
        r.ch.observe_ef(c0);
        r.ch.observe_ef(c2);
        if (pow_bits > 0) r.P->initial_sumcheck.pow_witnesses.push_back(grind_all(c, {&r.ch}, pow_bits)[0]);
        return r.ch.sample_ef();
      }));
    }
  }
}

// Leaf hashes and the tree above them for each proof's commitment (one shape),
// timed separately; the last tree kernel writes each root to its slot's next
// root entry.
void commit_trees(whirgpu_ctx* c, const std::vector<whirgpu_prover_data*>& pds) {
  const int B = (int)pds.size();
  const whirgpu_prover_data* p0 = pds[0];
  const int lw = (int)c->params.leaf_hash_width;
  const int rate = lw - 8;
  std::vector<TreeArgs> ta(B);
  for (int b = 0; b < B; b++) {
    whirgpu_prover_data* pd = pds[b];
    pd->root_slot = pd->sl->root_next;
    pd->sl->root_next = (pd->sl->root_next + 1) & 7;
    ta[b] = TreeArgs{pd->rows.p, pd->tree.p, pd->sl->d_root + 8 * pd->root_slot};
  }
  const bool keccak = c->params.mmcs == WHIRGPU_MMCS_KECCAK;
  {
    Span sp(c, ST_LEAVES, keccak ? 0.0 : (double)p0->height * (double)((p0->row_len + rate - 1) / rate) * B);
    KCHK(keccak ? keccak_leaves(ta.data(), B, p0->row_len, p0->height, c->st)
                : merkle_leaves(ta.data(), B, p0->row_len, p0->height, lw, c->d_c16, c->d_c24, c->st));
  }
  Span sp(c, ST_TREE, keccak ? 0.0 : (double)(p0->height - 1) * B);
  const int rc = keccak ? keccak_tree(ta.data(), B, p0->height, c->st)
                        : merkle_tree(ta.data(), B, p0->height, c->c16, c->tune, c->st);
  if (rc < 0) KCHK(rc);
  for (whirgpu_prover_data* pd : pds) pd->root_on_host = rc == 0;
}

// The root: written to its slot's h_root entry by the tree's last kernel, or
// (single leaf) copied.  completed: the tree kernels are known to have finished.
Digest read_root(whirgpu_ctx* c, const whirgpu_prover_data* pd, bool completed = false) {
  Digest d;
  if (pd->root_on_host) {
    if (!completed) c->sync();
    memcpy(d.data(), pd->sl->h_root + 8 * pd->root_slot, sizeof d);
  } else {
    c->d2h(d.data(), pd->tree.p + (2 * pd->height - 2) * 8, 8);
  }
  return d;
}

void observe_digest(Challenger& ch, const Digest& d) { ch.observe_digest(d); }

// Commitment of round r's folded polynomial for every proof (prover/mod.rs:
// 158-190): 2^(m - kn + R) rows of 2^kn extension elements and their tree.
// The commitments go to each run's `next`.
void commit_round(whirgpu_ctx* c, const std::vector<Run*>& R, int round) {
  const int B = (int)R.size();
  const auto& cfg = c->cfg;
  const int m = cfg.num_variables - cfg.params.folding_factor.total_number(round);
  const int kn = cfg.ff(round + 1);
  const int Rl = ilog2u(cfg.inv_rate(round));
  const u64 H = 1ull << (m - kn + Rl), W = 1ull << kn;
  const bool two = lde_two_pass(m, kn, true, c->tune);
  std::vector<Buf<ef>> tmps(B);
  std::vector<LdeArgs<ef>> la(B);
  std::vector<whirgpu_prover_data*> pds(B);
  for (int b = 0; b < B; b++) {
    auto* pd = new whirgpu_prover_data();
    R[b]->next = pd;   // destroyed by the caller if anything below throws
    pd->ctx = c;
    pd->sl = R[b]->sl;
    pd->rows = c->alloc<u32>(H * W * 4);
    pd->tree = c->alloc<u32>((2 * H - 1) * 8);
    pd->height = H;
    pd->row_len = (int)(4 * W);
    pd->ext = true;
    if (two) tmps[b] = c->alloc<ef>(H * W);
    la[b] = LdeArgs<ef>{R[b]->s.e.p, tmps[b].p, reinterpret_cast<ef*>(pd->rows.p)};
    pds[b] = pd;
  }
  {
    Span sp(c, ST_LDE_EXT, (16.0 * (double)(1ull << m) + 16.0 * (double)(H * W)) * B);
    KCHK(lde_ext(la.data(), B, m, kn, Rl, c->tw, c->tune, c->st));
  }
  commit_trees(c, pds);
  for (auto& t : tmps)
    if (t.p) c->release(t);   // reuse is ordered on the same stream
}

// Opens `idx` on one proof's committed matrix (host-driven rounds).
void parse_openings(const whirgpu_prover_data* pd, const u32* h, int nq, std::vector<whir::QueryOpening>* out,
                    std::vector<std::vector<ef>>* rows_ef);

void open_rows(whirgpu_ctx* c, Slot* sl, const whirgpu_prover_data* pd, const std::vector<u64>& idx,
               std::vector<whir::QueryOpening>* out, std::vector<std::vector<ef>>* rows_ef) {
  const int nq = (int)idx.size();
  if (nq == 0) return;
  const size_t per = (size_t)pd->row_len + 8ull * ilog2u(pd->height);
  if (per * nq > c->open_cap || (size_t)nq > Slot::IDX_CAP) bad("too many queries for the opening buffer");
  // h_idx / h_open are free: every opening ends with a synchronization.
  memcpy(sl->h_idx, idx.data(), idx.size() * sizeof(u64));
  const OpenArgs oa{pd->rows.p, pd->tree.p, sl->d_idx, nullptr, sl->d_open};
  KCHK(merkle_open(&oa, 1, pd->row_len, pd->height, nq, c->st));
  c->sync();
  parse_openings(pd, sl->h_open, nq, out, rows_ef);
}

// The openings of nq queries written to h by an opening kernel (after a
// synchronization), as proof entries and rows of extension elements.
void parse_openings(const whirgpu_prover_data* pd, const u32* h, int nq, std::vector<whir::QueryOpening>* out,
                    std::vector<std::vector<ef>>* rows_ef) {
  const int depth = ilog2u(pd->height);
  const size_t per = (size_t)pd->row_len + 8ull * depth;
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

// Whether a round's steps after its commitment fit the device transcript's
// buffers (otherwise the host drives them, one proof at a time).
bool device_round_fits(const whirgpu_ctx* c, const whir::RoundConfig& rp, int m, const whirgpu_prover_data* prev) {
  const int lo = std::min(8, m), hi = m - lo;
  if (c->params.mmcs != WHIRGPU_MMCS_POSEIDON2) return false;   // the device transcript is the duplex challenger's
  if (!prev || m < 1 || m > RT_MAX_VARS) return false;
  if (rp.ood_samples > RT_MAX_OOD || rp.num_queries < 1 || rp.num_queries > RT_MAX_QUERIES) return false;
  if ((1ull << hi) + (1ull << lo) > c->tab_cap) return false;
  const size_t per = (size_t)prev->row_len + 8ull * ilog2u(prev->height);
  return per * (size_t)rp.num_queries <= c->open_cap / 2;   // the second half may hold the final round's
}

// A round's steps from its commitment's root to the end of its sumcheck
// (prover/mod.rs:192-303) with the transcript on the device, for every proof
// of the group: the OOD samples and their answers, the proof of work, the
// STIR queries and their openings, gamma, the weight updates and the sumcheck
// run as one stream of launches, each reading each proof's challenger from
// the previous one, and one event at the end.  The caller may enqueue the
// next round's commitment (which needs none of the host's results) before
// device_round_replay waits for that event; the host then replays every step
// on each proof's challenger and requires each challenge, answer and query
// index to equal the device's, so the proof and the challenger state are
// those of the host-driven steps.
struct RoundLaunch {
  ScLaunch la;
  size_t grind_rec = SIZE_MAX;
};

RoundLaunch device_round_launch(whirgpu_ctx* c, const std::vector<Run*>& R, int round, int kn) {
  const int B = (int)R.size();
  const auto& cfg = c->cfg;
  const auto& rp = cfg.round_parameters[round];
  const int m = R[0]->s.m;
  const int n_ood = rp.ood_samples;
  const int lo = std::min(8, m), hi = m - lo;
  std::vector<const DevChallenger*> cin(B);
  std::vector<const u32*> obs(B);
  std::vector<ef*> log_obs(B, nullptr);
  std::vector<const ef*> obs_sum(B, nullptr);   // the last answer as its eq-weighted sum's block partials
  int obs_nb = 0;
  int n_obs = 8;
  int ri = 0;
  auto rch = [&](int b) { return R[b]->sl->d_rchal + ri; };
  for (int b = 0; b < B; b++) {
    upload_challenger(R[b]->sl, R[b]->ch.dx);
    cin[b] = R[b]->sl->d_chal_init;
    const whirgpu_prover_data* pd = R[b]->cur;
    obs[b] = pd->tree.p + (2 * pd->height - 2) * 8;   // the root
  }
  std::vector<OodArgs> oa(B);
  std::vector<DotArgs> da(B);
  if (n_ood > 0 && m <= OOD_SMALL_LOG && lo <= 8) {
    // every sample in one block per proof
    std::vector<OodSmallArgs> os(B);
    for (int b = 0; b < B; b++) {
      Slot* sl = R[b]->sl;
      os[b] = OodSmallArgs{StepTr{cin[b], rch(b), obs[b], n_obs, nullptr, nullptr, 0, nullptr}, sl->d_rlog->y, sl->d_rlog->ev,
                           sl->rb.pts,
                           sl->d_tabs, sl->tab_cap, R[b]->s.e.p, sl->rb.ev,
                           rp.pow_bits > 0 ? sl->rb.grind : nullptr};
    }
    KCHK(ood_small(os.data(), B, m, hi, lo, n_ood, c->d_c16, c->st));
    for (int b = 0; b < B; b++) {
      cin[b] = rch(b);
      obs[b] = reinterpret_cast<const u32*>(R[b]->sl->rb.ev + n_ood - 1);
      log_obs[b] = &R[b]->sl->d_rlog->ev[n_ood - 1];
    }
    n_obs = 4;
    ri ^= 1;
  }
  for (int i = 0; i < n_ood && !(m <= OOD_SMALL_LOG && lo <= 8); i++) {
    for (int b = 0; b < B; b++) {
      Slot* sl = R[b]->sl;
      // table slot i: the weight update below uses the same tables, scaled by gamma^i
      ef* tab = sl->d_tabs + (size_t)i * sl->tab_cap;
      oa[b] = OodArgs{StepTr{cin[b], rch(b), obs[b], n_obs, log_obs[b], obs_sum[b], obs_nb, nullptr}, &sl->d_rlog->y[i],
                      sl->rb.pts + (size_t)i * RT_MAX_VARS, tab, tab + (1ull << hi),
                      i == 0 && rp.pow_bits > 0 ? sl->rb.grind : nullptr};
      EqTables t{};
      t.hi[0] = tab;
      t.lo[0] = tab + (1ull << hi);
      da[b] = DotArgs{R[b]->s.e.p, t, sl->d_partial, nullptr};   // the next step reduces the partials
    }
    KCHK(ood_eq_tables(oa.data(), B, m, hi, lo, c->d_c16, c->st));
    {
      Span sp(c, ST_EQDOT, 16.0 * (double)(1ull << m) * B);
      KCHK(eq_dot_ext(da.data(), B, m, 0, lo, 1, dot_blocks(m - lo, 0), c->st));
    }
    for (int b = 0; b < B; b++) {
      cin[b] = rch(b);
      obs[b] = nullptr;
      obs_sum[b] = R[b]->sl->d_partial;
      log_obs[b] = &R[b]->sl->d_rlog->ev[i];
    }
    obs_nb = dot_blocks(m - lo, 0);
    n_obs = 4;
    ri ^= 1;
  }
  RoundLaunch rl;
  if (rp.pow_bits > 0) {
    std::vector<GrindArgs> ga(B);
    for (int b = 0; b < B; b++) {
      Slot* sl = R[b]->sl;
      // the first OOD step reset the grind slot; without OOD samples it is reset here
      if (n_ood == 0) HIPCHK(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(sl->rb.grind), (int)P, 1, c->st));
      memset(&ga[b], 0, sizeof ga[b]);
      ScTr& tr = ga[b].tr;
      tr.ch_in = cin[b];
      tr.ch_out = rch(b);
      tr.grind = sl->rb.grind;
      tr.pow_bits = rp.pow_bits;
      tr.obs = obs[b];
      tr.n_obs = n_obs;
      tr.log_obs = log_obs[b];
      tr.obs_sum = obs_sum[b];
      tr.obs_nb = obs_nb;
      tr.gbase = sl->d_gbase;
      ga[b].result = tr.grind;
    }
    Span sp(c, ST_GRIND, 0);
    KCHK(grind_range(c->d_c16, ga.data(), B, rp.pow_bits, TR_OBS_GRIND, 0, P, grind_blocks_for(c, rp.pow_bits),
                     c->tune, c->st));
    rl.grind_rec = sp.close();
    for (int b = 0; b < B; b++) {
      cin[b] = rch(b);
      obs[b] = nullptr;
      obs_sum[b] = nullptr;
      log_obs[b] = nullptr;
    }
    n_obs = 0;
    ri ^= 1;
  }
  {
    std::vector<QueryArgs> qa(B);
    for (int b = 0; b < B; b++)
      qa[b] = QueryArgs{StepTr{cin[b], rch(b), obs[b], n_obs, log_obs[b], obs_sum[b], obs_nb, nullptr}, R[b]->sl->rb.grind,
                        R[b]->sl->d_rlog,
                        R[b]->sl->rb};
    KCHK(round_queries(qa.data(), B, rp.pow_bits, rp.domain_size, cfg.ff(round), rp.num_queries,
                       rp.folded_domain_gen, n_ood, c->d_c16, c->st));
    for (int b = 0; b < B; b++) cin[b] = rch(b);
    ri ^= 1;
  }
  {
    // openings of the previous commitment into h_open, read after the round
    std::vector<OpenArgs> op(B);
    for (int b = 0; b < B; b++) {
      const whirgpu_prover_data* prev = R[b]->prev;
      op[b] = OpenArgs{prev->rows.p, prev->tree.p, R[b]->sl->rb.idx, R[b]->sl->rb.nq, R[b]->sl->d_open};
    }
    KCHK(merkle_open(op.data(), B, R[0]->prev->row_len, R[0]->prev->height, rp.num_queries, c->st));
  }
  // Constraint::combine_packed (src/whir/constraints/mod.rs:234-248): the OOD
  // points' eq (the tables of the OOD steps) weighted by gamma^i, then the
  // queries' select terms
  std::vector<CombineArgs> ca(B);
  for (int b0 = 0; b0 < n_ood; b0 += 4) {
    const int kk = std::min(4, n_ood - b0);
    for (int b = 0; b < B; b++) {
      Slot* sl = R[b]->sl;
      ca[b] = CombineArgs{R[b]->s.w.p, EqTables{}, sl->rb.gpow + b0};
      for (int k = 0; k < kk; k++) {
        ef* base = sl->d_tabs + (size_t)(b0 + k) * sl->tab_cap;
        ca[b].tabs.hi[k] = base;
        ca[b].tabs.lo[k] = base + (1ull << hi);
      }
    }
    Span sp(c, ST_COMBINE, 32.0 * (double)(1ull << m) * B);
    KCHK(eq_combine(ca.data(), B, 1ull << m, lo, kk, 1, c->st));
  }
  {
    const int ns = rp.num_queries;
    std::vector<Buf<ef>> As(B);
    std::vector<Buf<u32>> Bs(B);
    std::vector<SelectArgs> sa(B);
    for (int b = 0; b < B; b++) {
      Slot* sl = R[b]->sl;
      As[b] = c->alloc<ef>((size_t)ns << (m - m / 2));
      Bs[b] = c->alloc<u32>((size_t)ns << (m / 2));
      sa[b] = SelectArgs{R[b]->s.w.p, sl->rb.selz, sl->rb.gpow + n_ood, sl->rb.nq, As[b].p, Bs[b].p};
    }
    {
      Span sp(c, ST_COMBINE, 32.0 * (double)(1ull << m) * B);
      KCHK(select_combine(sa.data(), B, m, ns, c->st));
    }
    for (int b = 0; b < B; b++) {   // reuse is ordered on the same stream
      c->release(As[b]);
      c->release(Bs[b]);
    }
  }
  rl.la = sumcheck_launch(c, R, kn, rp.folding_pow_bits, cin);
  HIPCHK(hipEventRecord(c->ev_round, c->st));
  return rl;
}

// After the round launched by device_round_launch has completed (not what
// was enqueued after it): replays one proof's round on its host challenger.
void device_round_replay(whirgpu_ctx* c, Run& r, int round, const RoundLaunch& rl) {
  const auto& cfg = c->cfg;
  const auto& rp = cfg.round_parameters[round];
  const int n_ood = rp.ood_samples;
  whir::WhirRoundProof& RP = r.P->rounds[round];
  RP.commitment = read_root(c, r.cur, true);
  observe_digest(r.ch, RP.commitment);
  const RoundLog& L = *r.sl->h_rlog;
  for (int i = 0; i < n_ood; i++) {
    const ef y = r.ch.sample_ef();
    if (!ef_eq(y, L.y[i])) bad("device round transcript diverged (OOD challenge)", WHIRGPU_BAD_STATE);
    r.ch.observe_ef(L.ev[i]);
    RP.ood_answers.push_back(L.ev[i]);
  }
  if (rp.pow_bits > 0) {
    if (L.err) bad("device round transcript: no valid proof-of-work witness", WHIRGPU_GRIND_FAILED);
    if (!r.ch.check_witness(rp.pow_bits, L.witness))
      bad("device round transcript: witness rejected", WHIRGPU_GRIND_FAILED);
    RP.pow_witness = L.witness;
    if (rl.grind_rec != SIZE_MAX) c->recs[rl.grind_rec].work += (double)from_monty(L.witness) + 1;
  }
  r.ch.sample();
  const auto idx = whir::get_challenge_stir_queries(rp.domain_size, cfg.ff(round), rp.num_queries, r.ch);
  if (idx.size() != L.nq) bad("device round transcript diverged (query count)", WHIRGPU_BAD_STATE);
  for (size_t q = 0; q < idx.size(); q++)
    if (idx[q] != L.idx[q]) bad("device round transcript diverged (query index)", WHIRGPU_BAD_STATE);
  std::vector<std::vector<ef>> rows;
  parse_openings(r.prev, r.sl->h_open, (int)idx.size(), &RP.queries, &rows);
  const ef gamma = r.ch.sample_ef();
  if (!ef_eq(gamma, L.gamma)) bad("device round transcript diverged (gamma)", WHIRGPU_BAD_STATE);
  const auto gp = whir::powers(gamma, (size_t)n_ood + idx.size());
  for (int i = 0; i < n_ood; i++) r.s.sum = ef_add(r.s.sum, ef_mul(gp[i], L.ev[i]));
  for (size_t j = 0; j < idx.size(); j++)
    r.s.sum = ef_add(r.s.sum, ef_mul(gp[n_ood + j], whir::mle_host(rows[j], r.fold_rand)));
  r.fold_rand = sumcheck_replay(c, r, rl.la, RP.sumcheck);
}

// Whether the final round's steps fit the device transcript: the device
// observes the final polynomial word by word, so it must be small.  Its
// openings use the second half of the opening buffer.
bool device_final_fits(const whirgpu_ctx* c, const whir::WhirConfig& cfg, int m, const whirgpu_prover_data* prev) {
  if (c->params.mmcs != WHIRGPU_MMCS_POSEIDON2) return false;
  if (!prev || m < 0 || m > Slot::FINAL_MAX_LOG) return false;
  if (cfg.final_queries < 1 || cfg.final_queries > RT_MAX_QUERIES) return false;
  if (cfg.final_sumcheck_rounds > Slot::SC_LOG_CAP / 2) return false;
  const size_t per = (size_t)prev->row_len + 8ull * ilog2u(prev->height);
  return per * (size_t)cfg.final_queries <= c->open_cap / 2;
}

struct FinalLaunch {
  ScLaunch la;
  size_t grind_rec = SIZE_MAX;
  size_t nf = 0;
};

// The final round (Prover::final_round, prover/mod.rs:330-417) with the
// transcript on the device, for every proof of the group: observe the final
// polynomial, the proof of work, the final STIR queries and their openings,
// and the final sumcheck, as one stream of launches.  Proof b starts from the
// device challenger cin[b] (the last round's, so it can be enqueued before
// that round is replayed) or, with cin empty, from its host challenger.  Its
// log, openings and sumcheck log use their own slots (RoundLog 1, the opening
// buffer's second half, ScLog entries from SC_LOG_CAP / 2), apart from the
// last round's.
// grind_clear: the round's grind slot is known to hold p (a device round's
// query step, which sets it back after reading its witness, ran last).
FinalLaunch device_final_launch(whirgpu_ctx* c, const std::vector<Run*>& R, int round,
                                std::vector<const DevChallenger*> cin, bool grind_clear) {
  const int B = (int)R.size();
  const auto& cfg = c->cfg;
  FinalLaunch fl;
  const int m = R[0]->s.m;
  fl.nf = 1ull << m;
  const int pow = cfg.final_pow_bits;
  const auto fc = cfg.final_round_config();
  if (cin.empty()) {
    for (Run* r : R) {
      upload_challenger(r->sl, r->ch.dx);
      cin.push_back(r->sl->d_chal_init);
    }
  }
  int ri = 0;
  auto rch = [&](int b) { return R[b]->sl->d_rchal + ri; };
  // the step that observes the final polynomial also copies it to the slot's
  // small buffer (host-visible), before the final sumcheck folds it
  if (fl.nf > Slot::SMALL_CAP) bad("final polynomial larger than the small buffer", WHIRGPU_BAD_STATE);
  int n_obs = (int)(4 * fl.nf);
  bool observe_final = true;
  if (pow > 0) {
    std::vector<GrindArgs> ga(B);
    for (int b = 0; b < B; b++) {
      Slot* sl = R[b]->sl;
      if (!grind_clear) HIPCHK(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(sl->rb.grind), (int)P, 1, c->st));
      memset(&ga[b], 0, sizeof ga[b]);
      ScTr& tr = ga[b].tr;
      tr.ch_in = cin[b];
      tr.ch_out = rch(b);
      tr.grind = sl->rb.grind;
      tr.pow_bits = pow;
      tr.obs = reinterpret_cast<const u32*>(R[b]->s.e.p);
      tr.n_obs = n_obs;
      tr.obs_copy = reinterpret_cast<u32*>(sl->d_small);
      tr.gbase = sl->d_gbase;
      ga[b].result = tr.grind;
    }
    Span sp(c, ST_GRIND, 0);
    KCHK(grind_range(c->d_c16, ga.data(), B, pow, TR_OBS_GRIND, 0, P, grind_blocks_for(c, pow), c->tune, c->st));
    fl.grind_rec = sp.close();
    for (int b = 0; b < B; b++) cin[b] = rch(b);
    ri ^= 1;
    observe_final = false;
    n_obs = 0;
  }
  {
    std::vector<QueryArgs> qa(B);
    for (int b = 0; b < B; b++) {
      const u32* obs = observe_final ? reinterpret_cast<const u32*>(R[b]->s.e.p) : nullptr;
      qa[b] = QueryArgs{StepTr{cin[b], rch(b), obs, n_obs, nullptr, nullptr, 0,
                               observe_final ? reinterpret_cast<u32*>(R[b]->sl->d_small) : nullptr},
                        R[b]->sl->rb.grind,
                        R[b]->sl->d_rlog + 1,
                        R[b]->sl->rb};
    }
    KCHK(round_queries(qa.data(), B, pow, fc.domain_size, cfg.ff(round), cfg.final_queries, 1u, 0, c->d_c16, c->st,
                       true));
    for (int b = 0; b < B; b++) cin[b] = rch(b);
    ri ^= 1;
  }
  {
    std::vector<OpenArgs> op(B);
    for (int b = 0; b < B; b++) {
      const whirgpu_prover_data* prev = R[b]->prev;
      op[b] = OpenArgs{prev->rows.p, prev->tree.p, R[b]->sl->rb.idx, R[b]->sl->rb.nq,
                       R[b]->sl->d_open + c->open_cap / 2};
    }
    KCHK(merkle_open(op.data(), B, R[0]->prev->row_len, R[0]->prev->height, cfg.final_queries, c->st));
  }
  if (cfg.final_sumcheck_rounds > 0)
    fl.la = sumcheck_launch(c, R, cfg.final_sumcheck_rounds, cfg.final_folding_pow_bits, cin, Slot::SC_LOG_CAP / 2);
  return fl;
}

// After everything enqueued has completed: replays one proof's final round on
// its host challenger.
void device_final_replay(whirgpu_ctx* c, Run& r, int round, const FinalLaunch& fl) {
  const auto& cfg = c->cfg;
  const int pow = cfg.final_pow_bits;
  const auto fc = cfg.final_round_config();
  whir::WhirProof& Pf = *r.P;
  Pf.final_poly.assign(r.sl->h_small, r.sl->h_small + fl.nf);
  Pf.has_final_poly = true;
  for (const ef& e : Pf.final_poly) r.ch.observe_ef(e);
  const RoundLog& L = r.sl->h_rlog[1];
  if (pow > 0) {
    if (L.err) bad("device final round: no valid proof-of-work witness", WHIRGPU_GRIND_FAILED);
    if (!r.ch.check_witness(pow, L.witness)) bad("device final round: witness rejected", WHIRGPU_GRIND_FAILED);
    Pf.final_pow_witness = L.witness;
    if (fl.grind_rec != SIZE_MAX) c->recs[fl.grind_rec].work += (double)from_monty(L.witness) + 1;
  }
  const auto idx = whir::get_challenge_stir_queries(fc.domain_size, cfg.ff(round), cfg.final_queries, r.ch);
  if (idx.size() != L.nq) bad("device final round diverged (query count)", WHIRGPU_BAD_STATE);
  for (size_t q = 0; q < idx.size(); q++)
    if (idx[q] != L.idx[q]) bad("device final round diverged (query index)", WHIRGPU_BAD_STATE);
  std::vector<std::vector<ef>> rows;
  parse_openings(r.prev, r.sl->h_open + c->open_cap / 2, (int)idx.size(), &Pf.final_queries, &rows);
  if (cfg.final_sumcheck_rounds > 0) {
    Pf.has_final_sumcheck = true;
    r.fold_rand = sumcheck_replay(c, r, fl.la, Pf.final_sumcheck);
  }
}

// A WHIR round's steps driven by the host, for one proof (shapes beyond the
// device transcript's buffers): OOD samples, proof of work, STIR queries,
// gamma, the weight updates and the round's sumcheck.
// `rounds` sumcheck rounds of one proof into sd: on the device transcript for
// the Poseidon2 challenger; round by round from the host (coefficients on the
// device, observe_and_sample on the host, proof/mod.rs:163-193) for the
// Keccak challenger.
void run_sumcheck(whirgpu_ctx* c, Run& r, int rounds, int pow_bits, whir::SumcheckData& sd) {
  if (!r.ch.keccak) {
    device_sumcheck(c, {&r}, rounds, pow_bits, [&](Run&) -> whir::SumcheckData& { return sd; });
    return;
  }
  r.fold_rand.clear();
  for (int i = 0; i < rounds; i++) {
    r.fold_rand.push_back(sumcheck_round(c, r.sl, r.s, [&](const ef& c0, const ef& c2) {
      sd.polynomial_evaluations.push_back({c0, c2});
      r.ch.observe_ef(c0);
      r.ch.observe_ef(c2);
      if (pow_bits > 0) sd.pow_witnesses.push_back(grind_all(c, {&r.ch}, pow_bits)[0]);
      return r.ch.sample_ef();
    }));
  }
}

void host_round(whirgpu_ctx* c, Run& r, int round) {
  const auto& cfg = c->cfg;
  const auto& rp = cfg.round_parameters[round];
  const int m = r.s.m;
  const int kn = cfg.ff(round + 1);
  whir::WhirRoundProof& RP = r.P->rounds[round];
  RP.commitment = read_root(c, r.cur);
  observe_digest(r.ch, RP.commitment);
  // OOD samples (prover/mod.rs:192-205)
  std::vector<std::vector<ef>> ood_pts;
  for (int i = 0; i < rp.ood_samples; i++) {
    auto pt = whir::expand_from_univariate(r.ch.sample_ef(), m);
    const ef ev = ext_mle(c, r.sl, r.s.e.p, m, pt);
    r.ch.observe_ef(ev);
    RP.ood_answers.push_back(ev);
    ood_pts.push_back(pt);
  }
  if (rp.pow_bits > 0) RP.pow_witness = grind_all(c, {&r.ch}, rp.pow_bits)[0];
  r.ch.sample();
  // STIR queries on the previous commitment (prover/mod.rs:229-303)
  auto idx = whir::get_challenge_stir_queries(rp.domain_size, cfg.ff(round), rp.num_queries, r.ch);
  std::vector<std::vector<ef>> rows;
  open_rows(c, r.sl, r.prev, idx, &RP.queries, &rows);
  std::vector<u32> sel_z;
  std::vector<ef> sel_v;
  for (size_t q = 0; q < idx.size(); q++) {
    sel_z.push_back(kb::pow(rp.folded_domain_gen, idx[q]));
    sel_v.push_back(whir::mle_host(rows[q], r.fold_rand));
  }
  // Constraint::combine_packed (src/whir/constraints/mod.rs:234-248)
  const ef gamma = r.ch.sample_ef();
  const auto gp = whir::powers(gamma, ood_pts.size() + sel_z.size());
  std::vector<ef> eq_scales(gp.begin(), gp.begin() + ood_pts.size());
  if (!ood_pts.empty()) add_eq_weights(c, {&r}, {ood_pts}, {eq_scales}, true);
  for (size_t i = 0; i < ood_pts.size(); i++) r.s.sum = ef_add(r.s.sum, ef_mul(gp[i], RP.ood_answers[i]));
  if (!sel_z.empty()) {
    const size_t ns = sel_z.size();
    std::vector<ef> g(gp.begin() + ood_pts.size(), gp.end());
    // the points and weights in one staged upload: weights first (16-byte
    // aligned), then the points
    auto gs = c->stage<ef>(ns + (ns + 3) / 4);
    u32* zh = reinterpret_cast<u32*>(gs.first + ns);
    const u32* zd = reinterpret_cast<const u32*>(gs.second + ns);
    memcpy(gs.first, g.data(), ns * sizeof(ef));
    memcpy(zh, sel_z.data(), ns * sizeof(u32));
    c->flush();
    const std::pair<const u32*, const ef*> zs_dev{zd, gs.second};
    // sized for the round's query count (ns <= num_queries after
    // deduplication), so every proof of this shape reuses one pool pair
    const int lo = m / 2, hi = m - lo;
    const size_t nmax = std::max<size_t>(ns, (size_t)rp.num_queries);
    Buf<ef> A = c->alloc<ef>(nmax << hi);
    Buf<u32> Bt = c->alloc<u32>(nmax << lo);
    const SelectArgs sa{r.s.w.p, zs_dev.first, zs_dev.second, nullptr, A.p, Bt.p};
    {
      Span sp(c, ST_COMBINE, 32.0 * (double)(1ull << m));
      KCHK(select_combine(&sa, 1, m, (int)ns, c->st));
    }
    c->release(A);   // reuse is ordered on the same stream
    c->release(Bt);
    for (size_t j = 0; j < ns; j++) r.s.sum = ef_add(r.s.sum, ef_mul(g[j], sel_v[j]));
  }
  run_sumcheck(c, r, kn, rp.folding_pow_bits, RP.sumcheck);
}

// The final round driven by the host, for one proof (final_round, prover/mod.rs:330-417).
void host_final(whirgpu_ctx* c, Run& r, int round) {
  const auto& cfg = c->cfg;
  whir::WhirProof& Pf = *r.P;
  Pf.final_poly.resize(1ull << r.s.m);
  c->d2h(Pf.final_poly.data(), r.s.e.p, Pf.final_poly.size());
  Pf.has_final_poly = true;
  for (const ef& e : Pf.final_poly) r.ch.observe_ef(e);
  if (cfg.final_pow_bits > 0) Pf.final_pow_witness = grind_all(c, {&r.ch}, cfg.final_pow_bits)[0];
  const auto fc = cfg.final_round_config();
  const auto idx = whir::get_challenge_stir_queries(fc.domain_size, cfg.ff(round), cfg.final_queries, r.ch);
  std::vector<std::vector<ef>> rows;
  open_rows(c, r.sl, r.prev, idx, &Pf.final_queries, &rows);
  if (cfg.final_sumcheck_rounds > 0) {
    Pf.has_final_sumcheck = true;
    run_sumcheck(c, r, cfg.final_sumcheck_rounds, cfg.final_folding_pow_bits, Pf.final_sumcheck);
  }
}

// The commitment's root observation and OOD samples on the device transcript
// (commit_ood): per sample one step and the eq-weighted sums, a closing step,
// and one synchronization instead of one per sample; the host then replays
// the samples on its challengers (each y must equal the device's) and appends
// the points, answers and segment sums to the statements, as
// statement_evaluate would.  Returns false, having launched nothing, when the
// shapes do not fit; the caller then takes the host-driven path.
bool commit_ood_device(whirgpu_ctx* c, const std::vector<Run*>& R, const std::vector<whirgpu_prover_data*>& pds) {
  const auto& cfg = c->cfg;
  const int B = (int)R.size();
  const int n = cfg.num_variables;
  const int n_ood = cfg.commitment_ood_samples;
  if (n_ood < 1 || n_ood > RT_MAX_OOD || R[0]->ch.keccak || !pds[0]->root_on_host) return false;
  const int sb = R[0]->st->seg_bits;
  const int nrest = n - sb;
  if (sb > COOD_MAX_SB || n > RT_MAX_VARS || nrest < 1) return false;
  const int lo = std::min(8, nrest), hi = nrest - lo;
  const int nb = dot_blocks(nrest - lo, sb);
  const size_t S = 1ull << sb;
  if ((size_t)nb * S > c->partial_cap || S * n_ood + 2 * n_ood > Slot::SMALL_CAP) return false;
  for (int b = 0; b < B; b++)
    if (R[b]->st->seg_bits != sb || R[b]->st->n != n || (1ull << hi) + (1ull << lo) > R[b]->sl->tab_cap) return false;
  // host-visible layout in each slot's small buffer: the samples' segment
  // sums, then their answers, then their challenges y
  const size_t at_ev = S * n_ood, at_y = at_ev + n_ood;
  std::vector<const DevChallenger*> cin(B);
  for (int b = 0; b < B; b++) {
    upload_challenger(R[b]->sl, R[b]->ch.dx);
    cin[b] = R[b]->sl->d_chal_init;
  }
  int ri = 0;
  std::vector<CoodArgs> ca(B);
  std::vector<DotArgs> da(B);
  for (int i = 0; i <= n_ood; i++) {
    for (int b = 0; b < B; b++) {
      Slot* sl = R[b]->sl;
      const whirgpu_prover_data* pd = pds[b];
      DevChallenger* cout = sl->d_rchal + ri;
      CoodArgs& a = ca[b];
      a = CoodArgs{};
      a.st = StepTr{cin[b], cout, i == 0 ? pd->tree.p + (2 * pd->height - 2) * 8 : nullptr, i == 0 ? 8 : 0, nullptr,
                    nullptr, 0, nullptr};
      if (i > 0) {
        a.prev_part = sl->d_partial;
        a.prev_nb = nb;
        a.prev_z = sl->rb.pts + (size_t)(i - 1) * RT_MAX_VARS;
        a.prev_P = sl->d_small + S * (i - 1);
        a.prev_eval = sl->d_small + at_ev + (i - 1);
      }
      if (i < n_ood) {
        a.log_y = sl->d_small + at_y + i;
        a.z_out = sl->rb.pts + (size_t)i * RT_MAX_VARS;
        a.out_hi = sl->d_tabs;
        a.out_lo = sl->d_tabs + (1ull << hi);
      }
      cin[b] = cout;
      EqTables t{};
      t.hi[0] = sl->d_tabs;
      t.lo[0] = sl->d_tabs + (1ull << hi);
      da[b] = DotArgs{R[b]->st->d_poly, t, sl->d_partial, nullptr};   // reduced by the next step
    }
    KCHK(commit_ood(ca.data(), B, n, sb, hi, lo, i < n_ood ? 1 : 0, c->d_c16, c->st));
    ri ^= 1;
    if (i < n_ood) {
      Span sp(c, ST_EQDOT, 4.0 * (double)(1ull << n) * B);
      KCHK(eq_dot_base(da.data(), B, n, sb, lo, 1, nb, c->st));
    }
  }
  c->sync();
  for (int b = 0; b < B; b++) {
    Run& r = *R[b];
    const Digest root = read_root(c, pds[b], true);
    r.P->has_initial_commitment = true;
    r.P->initial_commitment = root;
    observe_digest(r.ch, root);
    const ef* small = r.sl->h_small;
    for (int i = 0; i < n_ood; i++) {
      const ef y = r.ch.sample_ef();
      if (!ef_eq(y, small[at_y + i])) bad("device transcript diverged from the host challenger", WHIRGPU_BAD_STATE);
      const ef ev = small[at_ev + i];
      r.st->points.push_back(whir::expand_from_univariate(y, n));
      r.st->evals.push_back(ev);
      r.st->partials.emplace_back(small + S * i, small + S * (i + 1));
      r.P->initial_ood_answers.push_back(ev);
      r.ch.observe_ef(ev);
    }
  }
  return true;
}

// CommitmentWriter::commit (src/whir/committer/writer.rs:56-109) for every
// proof of the group: the base LDE and Merkle tree of all polynomials in one
// launch set, the roots observed, and the commitment's OOD samples evaluated
// for all proofs at once.  Each run's commitment goes to run.prev.
void commit_group(whirgpu_ctx* c, const std::vector<Run*>& R) {
  const int B = (int)R.size();
  const auto& cfg = c->cfg;
  const int n = cfg.num_variables, k0 = cfg.ff(0), Rr = cfg.params.starting_log_inv_rate;
  const u64 H = 1ull << (n + Rr - k0), W = 1ull << k0;
  const bool two = lde_two_pass(n, k0, false, c->tune);
  std::vector<Buf<u32>> tmps(B);
  std::vector<LdeArgs<u32>> la(B);
  std::vector<whirgpu_prover_data*> pds(B);
  for (int b = 0; b < B; b++) {
    auto* pd = new whirgpu_prover_data();
    R[b]->prev = pd;   // destroyed by the caller if anything below throws
    pd->ctx = c;
    pd->sl = R[b]->sl;
    pd->rows = c->alloc<u32>(H * W);
    pd->tree = c->alloc<u32>((2 * H - 1) * 8);
    pd->height = H;
    pd->row_len = (int)W;
    pd->ext = false;
    if (two) tmps[b] = c->alloc<u32>(H * W);
    la[b] = LdeArgs<u32>{R[b]->st->d_poly, tmps[b].p, pd->rows.p};
    pds[b] = pd;
  }
  {
    Span sp(c, ST_LDE_BASE, (4.0 * (double)(1ull << n) + 4.0 * (double)(H * W)) * B);
    KCHK(lde_base(la.data(), B, n, k0, Rr, c->tw, c->tune, c->st));
  }
  commit_trees(c, pds);
  for (auto& t : tmps)
    if (t.p) c->release(t);
  if (commit_ood_device(c, R, pds)) return;
  if (pds[0]->root_on_host) c->sync();
  for (int b = 0; b < B; b++) {
    const Digest root = read_root(c, pds[b], true);
    R[b]->P->has_initial_commitment = true;
    R[b]->P->initial_commitment = root;
    observe_digest(R[b]->ch, root);
  }
  std::vector<whirgpu_statement*> sts(B);
  for (int b = 0; b < B; b++) sts[b] = R[b]->st;
  for (int i = 0; i < cfg.commitment_ood_samples; i++) {
    std::vector<std::vector<ef>> pts(B);
    for (int b = 0; b < B; b++) pts[b] = whir::expand_from_univariate(R[b]->ch.sample_ef(), n);
    const auto evs = statement_evaluate(c, sts, pts);
    for (int b = 0; b < B; b++) {
      R[b]->P->initial_ood_answers.push_back(evs[b]);
      R[b]->ch.observe_ef(evs[b]);
    }
  }
}

// Prover::prove (src/whir/prover/mod.rs:97-417) for every proof of the group,
// each starting from its commitment in run.prev (consumed).
void prove_group(whirgpu_ctx* c, const std::vector<Run*>& R) {
  const int B = (int)R.size();
  const auto& cfg = c->cfg;
  const int n = cfg.num_variables;
  for (Run* r : R) r->P->rounds.assign(cfg.n_rounds(), whir::WhirRoundProof());
  initial_sumcheck(c, R, cfg.ff(0), cfg.starting_folding_pow_bits);
  FinalLaunch fl;
  bool final_launched = false;
  bool last_device = false;   // the last WHIR round ran on the device transcript
  for (int round = 0; round <= cfg.n_rounds(); round++) {
    if (final_launched) {   // enqueued after the last round (its sumcheck already advanced s)
      c->sync();
      for (Run* r : R) device_final_replay(c, *r, round, fl);
      break;
    }
    const int m = n - cfg.params.folding_factor.total_number(round);
    for (Run* r : R)
      if (m != r->s.m) bad("internal: round size mismatch", WHIRGPU_BAD_STATE);
    if (round == cfg.n_rounds()) {
      if (device_final_fits(c, cfg, m, R[0]->prev)) {
        fl = device_final_launch(c, R, round, {}, last_device);
        c->sync();
        for (Run* r : R) device_final_replay(c, *r, round, fl);
      } else {
        for (Run* r : R) host_final(c, *r, round);
      }
      break;
    }
    const auto& rp = cfg.round_parameters[round];
    const int kn = cfg.ff(round + 1);
    if (!R[0]->next) commit_round(c, R, round);
    for (Run* r : R) {
      r->cur = r->next;
      r->next = nullptr;
    }
    if (device_round_fits(c, rp, m, R[0]->prev)) {
      const RoundLaunch rl = device_round_launch(c, R, round, kn);
      last_device = true;
      // the next commitment depends only on the folded polynomial; after
      // the last round, the final round continues from the device state
      if (round + 1 < cfg.n_rounds()) {
        commit_round(c, R, round + 1);
      } else if (kn <= Slot::SC_LOG_CAP / 2 && device_final_fits(c, cfg, m - kn, R[0]->cur)) {
        std::vector<Run*> RF = R;
        for (Run* r : RF) {   // the final round opens this round's commitment
          std::swap(r->prev, r->cur);
        }
        fl = device_final_launch(c, RF, round + 1, rl.la.out, true);
        for (Run* r : RF) std::swap(r->prev, r->cur);
        final_launched = true;
      }
      HIPCHK(hipEventSynchronize(c->ev_round));
      for (Run* r : R) device_round_replay(c, *r, round, rl);
    } else {
      for (Run* r : R) host_round(c, *r, round);
      last_device = false;
    }
    for (Run* r : R) {
      whirgpu_prover_data_destroy(r->prev);
      r->prev = r->cur;
      r->cur = nullptr;
    }
  }
  HIPCHK(hipStreamSynchronize(c->st));
  (void)B;
}
double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void use_device(const whirgpu_ctx* c) {
  if (c->device < 0 || !c->st) bad("host-only context has no device", WHIRGPU_BAD_STATE);
  HIPCHK(hipSetDevice(c->device));
}

Challenger load_ch(const whirgpu_ctx* c, const whirgpu_challenger* ch) {
  if (c->params.mmcs != WHIRGPU_MMCS_POSEIDON2) bad("this context's MMCS is Keccak: use the Keccak challenger");
  Challenger out;
  DuplexChallenger& d = out.dx;
  d.perm = &c->c16;
  memcpy(d.sponge, ch->sponge, sizeof d.sponge);
  memcpy(d.in, ch->input, sizeof d.in);
  memcpy(d.out, ch->output, sizeof d.out);
  d.in_len = (int)ch->input_len;
  d.out_len = (int)ch->output_len;
  if (d.in_len < 0 || d.in_len > 7 || d.out_len < 0 || d.out_len > 8) bad("invalid challenger state");
  return out;
}

// The Keccak challenger's state across the C ABI (its input bytes and the
// output bytes not yet popped).
Challenger load_kch(const whirgpu_ctx* c, const whirgpu_keccak_challenger* k) {
  if (c && c->params.mmcs != WHIRGPU_MMCS_KECCAK) bad("this context's MMCS is Poseidon2: use the duplex challenger");
  if (k->input_len > WHIRGPU_KECCAK_INPUT_CAP || k->output_len > 32) bad("invalid Keccak challenger state");
  Challenger ch;
  ch.keccak = true;
  ch.kc.in.assign(k->input, k->input + k->input_len);
  memcpy(ch.kc.out, k->output, sizeof ch.kc.out);
  ch.kc.out_len = (int)k->output_len;
  return ch;
}

void store_kch(const Challenger& ch, whirgpu_keccak_challenger* k) {
  if (ch.kc.in.size() > WHIRGPU_KECCAK_INPUT_CAP)
    bad("the Keccak challenger's input exceeds WHIRGPU_KECCAK_INPUT_CAP bytes", WHIRGPU_BAD_STATE);
  memcpy(k->input, ch.kc.in.data(), ch.kc.in.size());
  k->input_len = (u32)ch.kc.in.size();
  memcpy(k->output, ch.kc.out, sizeof k->output);
  k->output_len = (u32)ch.kc.out_len;
}

void store_ch(const Challenger& c, whirgpu_challenger* ch) {
  const DuplexChallenger& d = c.dx;
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
void check_lde_shapes(const whir::WhirConfig& cfg, const Tuning& t) {
  const int n = cfg.num_variables;
  int rc = lde_check(n, cfg.ff(0), cfg.params.starting_log_inv_rate, false, t);
  if (rc)
    bad("Reed-Solomon encoding of 2^" + std::to_string(n) + " coefficients in rows of 2^" + std::to_string(cfg.ff(0)) +
        " is not supported (code " + std::to_string(rc) + ")");
  for (int round = 0; round < cfg.n_rounds(); round++) {
    const int m = n - cfg.params.folding_factor.total_number(round);
    const int kn = cfg.ff(round + 1);
    const int Rl = ilog2u(cfg.inv_rate(round));
    rc = lde_check(m, kn, Rl, true, t);
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

namespace {

// The initial sumcheck of one constraint with challenges from `next` (no
// grinding): the split route for the first l0 rounds when n > l0, classic
// rounds otherwise; used by the kernel-level sumcheck entry point.
std::vector<ef> initial_sumcheck_scripted(whirgpu_ctx* c, Run& r, int l0, const ChallengeFn& next) {
  const whirgpu_statement* st = r.st;
  const std::vector<ef> ap(st->points.size(), ef_one());
  r.s.sum = ef_zero();
  for (size_t j = 0; j < st->points.size(); j++) r.s.sum = ef_add(r.s.sum, st->evals[j]);
  std::vector<ef> rs;
  if (l0 > 0 && st->seg_bits == l0 && st->n > l0) {
    auto P = st->partials;
    for (int i = 0; i < l0; i++) {
      ef c0, c2;
      svo_round_poly(st, P, ap, rs, i, l0, &c0, &c2);
      const ef x = next(c0, c2);
      r.s.sum = whir::extrapolate_012(c0, ef_sub(r.s.sum, c0), c2, x);
      rs.push_back(x);
      svo_fold_partials(P, x);
    }
    svo_finish(c, {&r}, l0, {ap}, {rs});
    return rs;
  }
  const int n = st->n;
  r.s.m = n;
  r.s.e = c->alloc<ef>(1ull << n);
  r.s.w = c->alloc<ef>(1ull << n);
  const ExtArgs xa{st->d_poly, r.s.e.p};
  KCHK(base_to_ext(&xa, 1, 1ull << n, c->st));
  add_eq_weights(c, {&r}, {st->points}, {ap}, false);
  for (int i = 0; i < l0; i++) rs.push_back(sumcheck_round(c, r.sl, r.s, next));
  return rs;
}

// Releases what a failed or finished group still holds.
void release_runs(whirgpu_ctx* c, std::vector<Run>& runs) {
  for (Run& r : runs) {
    for (whirgpu_prover_data** p : {&r.next, &r.cur, &r.prev}) {
      if (*p) whirgpu_prover_data_destroy(*p);
      *p = nullptr;
    }
    if (r.s.e.p) c->release(r.s.e);
    if (r.s.w.p) c->release(r.s.w);
  }
}

std::vector<ef> read_point(const uint32_t* p, uint32_t n) {
  std::vector<ef> z(n);
  for (uint32_t i = 0; i < n; i++) memcpy(z[i].c, p + 4 * i, 16);
  return z;
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
    if (params->mmcs != WHIRGPU_MMCS_POSEIDON2 && params->mmcs != WHIRGPU_MMCS_KECCAK)
      bad("mmcs must be WHIRGPU_MMCS_POSEIDON2 or WHIRGPU_MMCS_KECCAK");
    if (params->num_variables < 1 || params->num_variables > 28) bad("num_variables must be in [1, 28]");
    std::string err;
    if (!whir::WhirConfig::create((int)params->num_variables, p, &c->cfg, &err)) bad(err);
    check_lde_shapes(c->cfg, c->tune);
    load_perms(perms, &c->c16, &c->c24);
    if (device < 0) return;   // host-only context: schedule and transcript
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c->ev_round, hipEventDisableTiming));
    make_twiddles(c);
    const int n = (int)params->num_variables;
    c->tab_cap = (1ull << std::max(0, n - 8)) + 512;
    c->partial_cap = 1 << 15;
    c->open_cap = 1 << 20;
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    // One 256-lane block per CU: one wave on each of its four SIMDs.  More
    // waves add no throughput to the search loop (it is issue-bound) and
    // lengthen each pass over the candidates.
    c->grind_blocks = std::max(1, prop.multiProcessorCount);
    HIPCHK(hipMalloc(&c->d_c16, sizeof(P2C16)));
    HIPCHK(hipMemcpy(c->d_c16, &c->c16, sizeof(P2C16), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&c->d_c24, sizeof(P2C24)));
    HIPCHK(hipMemcpy(c->d_c24, &c->c24, sizeof(P2C24), hipMemcpyHostToDevice));
    const size_t gbytes = (size_t)whirgpu_ctx::GRIND_SLOTS * GROUP_MAX * sizeof(u32);
    HIPCHK(hipMalloc(&c->d_hgrind, gbytes));
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->h_hgrind), GROUP_MAX * sizeof(u32), hipHostMallocDefault));
    c->ring_cap = 1 << 20;
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->h_ring), c->ring_cap, hipHostMallocDefault));
    HIPCHK(hipMalloc(reinterpret_cast<void**>(&c->d_ring), c->ring_cap));
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->h_pts), sizeof(ef) * 32 * whirgpu_ctx::PTS_CAP,
                         hipHostMallocCoherent));
    HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void**>(&c->d_pts), c->h_pts, 0));
    c->fixed_bytes = 8192 * sizeof(u32) + sizeof(P2C16) + sizeof(P2C24) + gbytes + c->ring_cap;
    c->pin<u32>(1);
    ensure_slots(c, 1);
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
  for (auto& s : c->slots) slot_free(s.get());
  c->slots.clear();
  for (void* p : {(void*)c->d_tw, (void*)c->d_c16, (void*)c->d_c24, (void*)c->d_hgrind, (void*)c->d_ring})
    if (p) (void)hipFree(p);
  for (void* h : {(void*)c->h_pin, (void*)c->h_ring, (void*)c->h_hgrind, (void*)c->h_pts})
    if (h) (void)hipHostFree(h);
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
    Challenger d = load_ch(c, ch);
    d.observe_slice(values, n);
    store_ch(d, ch);
  });
}

int whirgpu_challenger_observe_domain_separator(const whirgpu_ctx* c, whirgpu_challenger* ch) {
  return guarded([&] {
    if (!c || !ch) bad("null argument");
    Challenger d = load_ch(c, ch);
    auto v = whir::domain_separator_pattern(c->cfg, (int)c->params.digest_elems);
    d.observe_slice(v.data(), v.size());
    store_ch(d, ch);
  });
}

int whirgpu_challenger_sample(const whirgpu_ctx* c, whirgpu_challenger* ch, uint32_t* out, size_t n) {
  return guarded([&] {
    if (!c || !ch || (!out && n)) bad("null argument");
    Challenger d = load_ch(c, ch);
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


int whirgpu_statement_evaluate(whirgpu_statement* s, const uint32_t* point, uint32_t* eval_out) {
  return guarded([&] {
    if (!s || !point || !eval_out) bad("null argument");
    use_device(s->ctx);
    const ef e = statement_evaluate(s->ctx, {s}, {read_point(point, (uint32_t)s->n)})[0];
    memcpy(eval_out, e.c, 16);
  });
}

int whirgpu_statement_evaluate_group(uint32_t count, whirgpu_statement* const* statements, const uint32_t* points,
                                     uint32_t* evals_out) {
  return guarded([&] {
    if (!statements || !points || !evals_out || count < 1 || count > GROUP_MAX) bad("bad argument");
    std::vector<whirgpu_statement*> sts(statements, statements + count);
    for (auto* s : sts)
      if (!s) bad("null statement");
    whirgpu_ctx* c = sts[0]->ctx;
    use_device(c);
    const uint32_t n = (uint32_t)sts[0]->n;
    std::vector<std::vector<ef>> zs;
    for (uint32_t b = 0; b < count; b++) zs.push_back(read_point(points + (size_t)4 * n * b, n));
    const auto ev = statement_evaluate(c, sts, zs);
    for (uint32_t b = 0; b < count; b++) memcpy(evals_out + 4 * b, ev[b].c, 16);
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

// Launch shapes (no reference counterpart).
int whirgpu_get_tuning(const whirgpu_ctx* c, whirgpu_tuning* out) {
  return guarded([&] {
    if (!c || !out) bad("null argument");
    const Tuning& t = c->tune;
    *out = whirgpu_tuning{t.lde_lp2[0], t.lde_lp2[1], t.lde_tb_log[0], t.lde_tb_log[1], t.tree_lane_log,
                          t.tree_s128_log, t.grind_wide_bits, t.grind_lock_bits, t.grind_blocks};
  });
}

int whirgpu_device_name(int device, char* out, size_t cap) {
  return guarded([&] {
    if (!out || cap == 0) bad("null argument");
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    snprintf(out, cap, "%s%s%s (%d CUs)", prop.name, prop.name[0] ? " " : "", prop.gcnArchName,
             prop.multiProcessorCount);
  });
}

int whirgpu_set_tuning(whirgpu_ctx* c, const whirgpu_tuning* in) {
  return guarded([&] {
    if (!c || !in) bad("null argument");
    Tuning t;
    t.lde_lp2[0] = in->lde_lp2_base;
    t.lde_lp2[1] = in->lde_lp2_ext;
    t.lde_tb_log[0] = in->lde_tb_log_base;
    t.lde_tb_log[1] = in->lde_tb_log_ext;
    t.tree_lane_log = in->tree_lane_log;
    t.tree_s128_log = in->tree_s128_log;
    t.grind_wide_bits = in->grind_wide_bits;
    t.grind_lock_bits = in->grind_lock_bits;
    t.grind_blocks = in->grind_blocks;
    if (!tuning_valid(t)) bad("tuning value out of range");
    check_lde_shapes(c->cfg, t);
    c->tune = t;
  });
}

// serde deserialization of WhirProof (src/whir/proof.rs:11-141).
int whirgpu_proof_deserialize(whirgpu_ctx* c, const uint8_t* bytes, size_t len, whirgpu_proof** out) {
  return guarded([&] {
    if (!c || (!bytes && len) || !out) bad("null argument");
    auto p = std::make_unique<whirgpu_proof>();
    p->ctx = c;
    std::string err;
    if (!whir::deserialize(bytes, len, &p->p, &err, c->params.mmcs == WHIRGPU_MMCS_POSEIDON2))
      bad(err, WHIRGPU_MALFORMED_PROOF);
    *out = p.release();
  });
}

// CommitmentReader::parse + Verifier::verify (committer/reader.rs:40-82,
// verifier/mod.rs:58-200) on the host, for the statement {points -> evals}.
}  // extern "C"

template <class Load, class Store>
int verify_one(const whirgpu_ctx* c, const whirgpu_proof* proof, uint32_t npoints, const uint32_t* points,
               const uint32_t* evals, Load&& load, Store&& store) {
  return guarded([&] {
    if (!c || !proof || (npoints && (!points || !evals))) bad("null argument");
    const int n = c->cfg.num_variables;
    std::vector<std::vector<ef>> pts(npoints, std::vector<ef>(n));
    std::vector<ef> evs(npoints);
    for (uint32_t k = 0; k < npoints; k++) {
      memcpy(pts[k].data(), points + (size_t)4 * n * k, sizeof(ef) * n);
      memcpy(&evs[k], evals + 4 * k, sizeof(ef));
    }
    Challenger d = load();
    whir::VerifierKey key{&c->cfg, &c->c16, &c->c24, (int)c->params.leaf_hash_width,
                          c->params.mmcs == WHIRGPU_MMCS_KECCAK};
    std::string why;
    if (whir::verify(key, proof->p, d, pts, evs, &why) != whir::VerifyFailure::None) bad(why, WHIRGPU_VERIFY_FAILED);
    store(d);
  });
}

extern "C" {

int whirgpu_verify(const whirgpu_ctx* c, const whirgpu_proof* proof, whirgpu_challenger* ch, uint32_t npoints,
                   const uint32_t* points, const uint32_t* evals) {
  return verify_one(
      c, proof, npoints, points, evals,
      [&] {
        if (!ch) bad("null argument");
        return load_ch(c, ch);
      },
      [&](const Challenger& d) { store_ch(d, ch); });
}

int whirgpu_verify_keccak(const whirgpu_ctx* c, const whirgpu_proof* proof, whirgpu_keccak_challenger* kch,
                          uint32_t npoints, const uint32_t* points, const uint32_t* evals) {
  return verify_one(
      c, proof, npoints, points, evals,
      [&] {
        if (!kch) bad("null argument");
        return load_kch(c, kch);
      },
      [&](const Challenger& d) { store_kch(d, kch); });
}

// The Keccak challenger (SerializingChallenger32<F, HashChallenger<u8,
// Keccak256Hash, 32>>) across the C ABI.
int whirgpu_keccak_challenger_new(whirgpu_keccak_challenger* k) {
  if (!k) return WHIRGPU_BAD_PARAMS;
  memset(k, 0, sizeof *k);
  return WHIRGPU_OK;
}

int whirgpu_keccak_challenger_observe(const whirgpu_ctx* c, whirgpu_keccak_challenger* k, const uint32_t* values,
                                      size_t n) {
  return guarded([&] {
    if (!k || (n && !values)) bad("null argument");
    Challenger d = load_kch(c, k);
    d.observe_slice(values, n);
    store_kch(d, k);
  });
}

int whirgpu_keccak_challenger_sample(const whirgpu_ctx* c, whirgpu_keccak_challenger* k, uint32_t* out, size_t n) {
  return guarded([&] {
    if (!k || (n && !out)) bad("null argument");
    Challenger d = load_kch(c, k);
    for (size_t i = 0; i < n; i++) out[i] = d.sample();
    store_kch(d, k);
  });
}

int whirgpu_keccak_challenger_observe_bytes(const whirgpu_ctx* c, whirgpu_keccak_challenger* k, const uint8_t* bytes,
                                            size_t n) {
  return guarded([&] {
    if (!k || (n && !bytes)) bad("null argument");
    Challenger d = load_kch(c, k);
    for (size_t i = 0; i < n; i++) d.kc.observe_byte(bytes[i]);
    store_kch(d, k);
  });
}

int whirgpu_keccak_challenger_sample_bits(const whirgpu_ctx* c, whirgpu_keccak_challenger* k, uint32_t bits,
                                          uint32_t* out) {
  return guarded([&] {
    if (!k || !out) bad("null argument");
    if (bits > 31) bad("bits must be at most 31");
    Challenger d = load_kch(c, k);
    *out = (uint32_t)d.sample_bits((int)bits);
    store_kch(d, k);
  });
}

int whirgpu_keccak_challenger_observe_domain_separator(const whirgpu_ctx* c, whirgpu_keccak_challenger* k) {
  return guarded([&] {
    if (!c || !k) bad("null argument");
    Challenger d = load_kch(c, k);
    auto v = whir::domain_separator_pattern(c->cfg, (int)c->params.digest_elems);
    d.observe_slice(v.data(), v.size());
    store_kch(d, k);
  });
}

int whirgpu_keccak_grind(whirgpu_ctx* c, whirgpu_keccak_challenger* k, uint32_t bits, uint32_t* witness_out) {
  return guarded([&] {
    if (!c || !k || !witness_out) bad("null argument");
    if (bits == 0 || bits > 30) bad("bits must be in [1, 30]");
    use_device(c);
    Challenger ch = load_kch(c, k);
    *witness_out = keccak_gpu_grind(c, ch, (int)bits);
    store_kch(ch, k);
  });
}

int whirgpu_keccak_f1600(uint64_t* state) {
  if (!state) return WHIRGPU_BAD_PARAMS;
  keccak_f1600(state);
  return WHIRGPU_OK;
}

int whirgpu_prover_data_destroy(whirgpu_prover_data* pd) {
  if (!pd) return WHIRGPU_OK;
  pd->ctx->release(pd->rows);
  pd->ctx->release(pd->tree);
  delete pd;
  return WHIRGPU_OK;
}

}  // extern "C"

// CommitmentWriter::commit (src/whir/committer/writer.rs:56-109): a group of
// one, from the challenger `load` gives; `store` writes it back on success.
template <class Load, class Store>
int commit_one(whirgpu_ctx* c, whirgpu_proof* proof, whirgpu_statement* st, whirgpu_prover_data** out, Load&& load,
               Store&& store) {
  std::vector<Run> runs(1);
  int rc = guarded([&] {
    if (!c || !proof || !st || !out) bad("null argument");
    if (st->ctx != c) bad("statement of another context");
    use_device(c);
    const double t0 = now_ms();
    Run& r = runs[0];
    r.st = st;
    r.P = &proof->p;
    r.ch = load();
    r.sl = c->slots[0].get();
    commit_group(c, {&r});
    store(r.ch);
    c->timings[0] = now_ms() - t0;
  });
  if (rc) {
    if (c) (void)hipStreamSynchronize(c->st);   // nothing in flight uses what is released below
    if (c) release_runs(c, runs);
    return rc;
  }
  *out = runs[0].prev;
  return WHIRGPU_OK;
}

// Prover::prove (src/whir/prover/mod.rs:97-417): a group of one; consumes the prover data.
template <class Load, class Store>
int prove_one(whirgpu_ctx* c, whirgpu_proof* proof, const whirgpu_statement* st, whirgpu_prover_data* pd0,
              Load&& load, Store&& store) {
  std::vector<Run> runs(1);
  runs[0].prev = pd0;
  int rc = guarded([&] {
    if (!c || !proof || !st || !pd0) bad("null argument");
    use_device(c);
    const double t0 = now_ms();
    Run& r = runs[0];
    r.st = const_cast<whirgpu_statement*>(st);
    r.P = &proof->p;
    r.ch = load();
    r.sl = pd0->sl ? pd0->sl : c->slots[0].get();
    prove_group(c, {&r});
    store(r.ch);
    c->timings[6] = now_ms() - t0;
  });
  if (c) {
    if (rc) (void)hipStreamSynchronize(c->st);
    release_runs(c, runs);
  }
  return rc;
}

extern "C" {

int whirgpu_commit(whirgpu_ctx* c, whirgpu_proof* proof, whirgpu_challenger* chs, whirgpu_statement* st,
                   whirgpu_prover_data** out) {
  return commit_one(
      c, proof, st, out,
      [&] {
        if (!chs) bad("null argument");
        return load_ch(c, chs);
      },
      [&](const Challenger& ch) { store_ch(ch, chs); });
}

int whirgpu_prove(whirgpu_ctx* c, whirgpu_proof* proof, whirgpu_challenger* chs, const whirgpu_statement* st,
                  whirgpu_prover_data* pd0) {
  return prove_one(
      c, proof, st, pd0,
      [&] {
        if (!chs) bad("null argument");
        return load_ch(c, chs);
      },
      [&](const Challenger& ch) { store_ch(ch, chs); });
}

// The Keccak variant (src/whir/mod.rs:236-383) of commit and prove.
int whirgpu_commit_keccak(whirgpu_ctx* c, whirgpu_proof* proof, whirgpu_keccak_challenger* kch, whirgpu_statement* st,
                          whirgpu_prover_data** out) {
  return commit_one(
      c, proof, st, out,
      [&] {
        if (!kch) bad("null argument");
        return load_kch(c, kch);
      },
      [&](const Challenger& ch) { store_kch(ch, kch); });
}

int whirgpu_prove_keccak(whirgpu_ctx* c, whirgpu_proof* proof, whirgpu_keccak_challenger* kch,
                         const whirgpu_statement* st, whirgpu_prover_data* pd0) {
  return prove_one(
      c, proof, st, pd0,
      [&] {
        if (!kch) bad("null argument");
        return load_kch(c, kch);
      },
      [&](const Challenger& ch) { store_kch(ch, kch); });
}

int whirgpu_prove_group(whirgpu_ctx* c, uint32_t count, whirgpu_statement* const* statements,
                        whirgpu_proof* const* proofs, whirgpu_challenger* challengers) {
  std::vector<Run> runs;
  int rc = guarded([&] {
    if (!c || !statements || !proofs || !challengers || count < 1 || count > GROUP_MAX) bad("bad argument");
    if (c->params.mmcs != WHIRGPU_MMCS_POSEIDON2)
      bad("group proving runs the duplex challenger's transcript on the device; this context's MMCS is Keccak");
    use_device(c);
    ensure_slots(c, (int)count);
    const double t0 = now_ms();
    runs.resize(count);
    std::vector<Run*> R;
    for (uint32_t b = 0; b < count; b++) {
      if (!statements[b] || !proofs[b]) bad("null statement or proof");
      if (statements[b]->ctx != c) bad("statement of another context");
      if (statements[b]->points.size() != statements[0]->points.size())
        bad("the statements of a group must have the same number of constraints");
      Run& r = runs[b];
      r.st = statements[b];
      r.P = &proofs[b]->p;
      r.ch = load_ch(c, &challengers[b]);
      r.sl = c->slots[b].get();
      R.push_back(&r);
    }
    commit_group(c, R);
    c->timings[0] = now_ms() - t0;
    prove_group(c, R);
    for (uint32_t b = 0; b < count; b++) store_ch(runs[b].ch, &challengers[b]);
    c->timings[6] = now_ms() - t0;
  });
  if (c) {
    if (rc) (void)hipStreamSynchronize(c->st);
    release_runs(c, runs);
  }
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
    const u64 out = 1ull << (m + log_blowup);
    Buf<u32> tmp;
    if (lde_two_pass((int)m, (int)k, false, c->tune)) tmp = c->alloc<u32>(out);
    const LdeArgs<u32> a{d_src, tmp.p, d_dst};
    KCHK(lde_base(&a, 1, (int)m, (int)k, (int)log_blowup, c->tw, c->tune, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (tmp.p) c->release(tmp);
  });
}

int whirgpu_lde_ext(whirgpu_ctx* c, const uint32_t* d_src, uint32_t m, uint32_t k, uint32_t log_blowup,
                    uint32_t* d_dst) {
  return guarded([&] {
    if (!c || !d_src || !d_dst || k > m || m > 28 || m - k + log_blowup > 24) bad("bad shape");
    use_device(c);
    const u64 out = 1ull << (m + log_blowup);
    Buf<ef> tmp;
    if (lde_two_pass((int)m, (int)k, true, c->tune)) tmp = c->alloc<ef>(out);
    const LdeArgs<ef> a{reinterpret_cast<const ef*>(d_src), tmp.p, reinterpret_cast<ef*>(d_dst)};
    KCHK(lde_ext(&a, 1, (int)m, (int)k, (int)log_blowup, c->tw, c->tune, c->st));
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
    const TreeArgs a{d_rows, d_tree, nullptr};
    const bool keccak = c->params.mmcs == WHIRGPU_MMCS_KECCAK;
    KCHK(keccak ? keccak_leaves(&a, 1, (int)row_len, height, c->st)
                : merkle_leaves(&a, 1, (int)row_len, height, (int)c->params.leaf_hash_width, c->d_c16, c->d_c24,
                                c->st));
    const int rc = keccak ? keccak_tree(&a, 1, height, c->st) : merkle_tree(&a, 1, height, c->c16, c->tune, c->st);
    if (rc < 0) KCHK(rc);
    c->d2h(root_out, d_tree + (2 * height - 2) * 8, 8);
  });
}

int whirgpu_grind(whirgpu_ctx* c, whirgpu_challenger* chs, uint32_t bits, uint32_t* witness_out) {
  return guarded([&] {
    if (!c || !chs || !witness_out) bad("null argument");
    if (bits == 0 || bits > 30) bad("bits must be in [1, 30]");
    use_device(c);
    Challenger ch = load_ch(c, chs);
    *witness_out = gpu_grind(c, {&ch.dx}, (int)bits)[0];
    store_ch(ch, chs);
  });
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
    Slot* sl = c->slots[0].get();
    base_partials(c, {d_f}, {sl}, (int)n, 0, {read_point(point, n)});
    ef e;
    read_small(c, sl, &e, 1);
    memcpy(out, e.c, 16);
  });
}

int whirgpu_mle_eval_ext(whirgpu_ctx* c, const uint32_t* d_f, uint32_t n, const uint32_t* point, uint32_t* out) {
  return guarded([&] {
    if (!c || !d_f || !point || !out || n > 30) bad("bad argument");
    if ((1ull << std::max<int>(0, (int)n - 8)) + 256 > c->tab_cap) bad("n too large for this context");
    use_device(c);
    const ef e = ext_mle(c, c->slots[0].get(), reinterpret_cast<const ef*>(d_f), (int)n, read_point(point, n));
    memcpy(out, e.c, 16);
  });
}

// Evaluation of a base-field multilinear polynomial without allocating (no
// counterpart in the reference: for callers that cannot allocate).  The same
// kernels as whirgpu_mle_eval_base, with the split eq tables and the block
// partial sums in the library's fixed device array `scratch` (its address on
// the current device comes from scratch_on_current_device), the point read
// from the caller's host memory, and the result written to device memory on
// the caller's stream.
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
    // the point goes down by value: a copy of the eq-table kernel's staging
    // would need memory of the library's own
    const PointArg a = point_arg(read_point(point, n));
    KCHK(eq_table(a, 0, hi, ef_one(), tab_hi, st));
    KCHK(eq_table(a, hi, lo, ef_one(), tab_lo, st));
    EqTables t{};
    t.hi[0] = tab_hi;
    t.lo[0] = tab_lo;
    const DotArgs da{d_f, t, partial, reinterpret_cast<ef*>(d_out)};
    KCHK(eq_dot_base(&da, 1, (int)n, 0, lo, 1, nb, st));
  });
}

int whirgpu_sumcheck_coeffs(whirgpu_ctx* c, const uint32_t* d_e, const uint32_t* d_w, uint32_t n, uint32_t* out) {
  return guarded([&] {
    if (!c || !d_e || !d_w || !out || n < 1 || n > 32) bad("bad argument");
    use_device(c);
    Slot* sl = c->slots[0].get();
    ScArgs a{};
    a.e = const_cast<ef*>(reinterpret_cast<const ef*>(d_e));
    a.w = const_cast<ef*>(reinterpret_cast<const ef*>(d_w));
    a.partial = sl->d_partial;
    a.c0c2 = sl->d_small;
    KCHK(sc_coeffs(&a, 1, 1ull << (n - 1), c->st));
    read_small(c, sl, reinterpret_cast<ef*>(out), 2);
  });
}

int whirgpu_sumcheck_fold(whirgpu_ctx* c, uint32_t* d_e, uint32_t* d_w, uint32_t n, const uint32_t* r, uint32_t* out) {
  return guarded([&] {
    if (!c || !d_e || !d_w || !r || n < 1 || n > 32) bad("bad argument");
    use_device(c);
    Slot* sl = c->slots[0].get();
    ScArgs a{};
    a.e = reinterpret_cast<ef*>(d_e);
    a.w = reinterpret_cast<ef*>(d_w);
    memcpy(a.r.c, r, 16);
    a.partial = sl->d_partial;
    a.c0c2 = sl->d_small;
    const u64 h = 1ull << (n - 1);
    KCHK(sc_fold(&a, 1, h, c->st));
    if (out && h >= 2) read_small(c, sl, reinterpret_cast<ef*>(out), 2);
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
  std::vector<Run> runs(1);
  int rc = guarded([&] {
    if (!c || !d_f || !point || !out_polys || !out_final || n < 1 || n > 30 || l0 > n) bad("bad argument");
    if (!scripted && !chs) bad("need scripted challenges or a challenger");
    use_device(c);
    Slot* sl = c->slots[0].get();
    ensure_tabs(c, sl, (int)n);
    whirgpu_statement st;
    st.ctx = c;
    st.n = (int)n;
    st.d_poly = const_cast<u32*>(d_f);
    st.seg_bits = (l0 > 0 && n > l0) ? (int)l0 : 0;
    statement_evaluate(c, {&st}, {read_point(point, n)});
    Run& r = runs[0];
    r.st = &st;
    r.sl = sl;
    Challenger ch;
    if (chs) ch = load_ch(c, chs);
    size_t round = 0;
    ChallengeFn next = [&](const ef& c0, const ef& c2) {
      memcpy(out_polys + 8 * round, c0.c, 16);
      memcpy(out_polys + 8 * round + 4, c2.c, 16);
      ef x;
      if (scripted) {
        memcpy(x.c, scripted + 4 * round, 16);
      } else {
        ch.observe_ef(c0);
        ch.observe_ef(c2);
        x = ch.sample_ef();
      }
      round++;
      return x;
    };
    initial_sumcheck_scripted(c, r, (int)l0, next);
    if (!scripted && r.s.m > 0) {
      // the remaining rounds on the device transcript (one synchronization),
      // replayed on the host challenger
      r.ch = ch;
      whir::SumcheckData sd;
      device_sumcheck(c, {&r}, r.s.m, 0, [&](Run&) -> whir::SumcheckData& { return sd; });
      for (const auto& pe : sd.polynomial_evaluations) {
        memcpy(out_polys + 8 * round, pe[0].c, 16);
        memcpy(out_polys + 8 * round + 4, pe[1].c, 16);
        round++;
      }
      ch = r.ch;
    }
    while (r.s.m > 0) sumcheck_round(c, sl, r.s, next);
    ef fin[2];
    c->d2h(&fin[0], r.s.e.p, 1);
    c->d2h(&fin[1], r.s.w.p, 1);
    memcpy(out_final, fin[0].c, 16);
    memcpy(out_final + 4, fin[1].c, 16);
    if (!ef_eq(ef_mul(fin[0], fin[1]), r.s.sum)) bad("internal: final sumcheck claim mismatch", WHIRGPU_BAD_STATE);
    if (chs) store_ch(ch, chs);
  });
  if (c) release_runs(c, runs);
  return rc;
}

}  // extern "C"
