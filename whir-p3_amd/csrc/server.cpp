// Batched and served proving on one device: a FIFO of jobs and a pool of host
// worker threads.  A free worker takes the job at the head of the queue and,
// with it, up to group_max - 1 more queued jobs of the same shape (number of
// variables, number of constraints, input on host or device), and proves
// them as one group (whirgpu_prove_group: every device step one launch for
// the whole group).  Jobs submitted together are queued largest first, so
// the small sizes fill in at the end, and a worker prefers jobs of a size it
// already holds a context for (pick_job), so repeated batches reuse contexts
// and pooled buffers instead of allocating.  (Capping a group at its share of the
// queued work, to spread the largest sizes over idle workers, was measured
// and was slower: large groups amortize more than a balanced tail gains.)  Each worker keeps its own contexts (one per number of
// variables, each with its own HIP stream and buffer pool), so groups of
// different workers overlap on the device: while one worker waits for a
// round's transcript values another worker's kernels run.  A group that fails
// is retried job by job, so one bad job does not fail the others.  With the
// self-check on, each proof is verified on the host (whirgpu_verify, from a
// fresh transcript) before its bytes are published; a rejected proof fails
// its job with WHIRGPU_VERIFY_FAILED.
//
// Admission by device memory.  A context keeps its buffers between jobs, so
// what the server holds on the device is what the workers' contexts hold, not
// what the running jobs need.  The server tracks both: `held`, the bytes of
// every worker's contexts as the library reports them after each job, and
// `reserved`, the growth each running job may still add (the peak expected for
// a context of its size minus what the worker's context of that size already
// holds).  A free worker takes the job at the head of the queue when
// held + reserved + growth fits in the budget and reserved + growth fits in
// the device memory free at that moment.  When it does not fit, idle workers
// release their contexts of other sizes (a worker that has proved this size
// before needs little or nothing more).  With nothing running and nothing left
// to release the job is admitted anyway, so an oversized job still completes.
//
// This file uses the prover only through the C ABI and internal.h (no HIP
// calls of its own), so its threading is checked on the host, under
// ThreadSanitizer, against a stand-in library (tests/server_standin.cpp).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/whir_gpu.h"
#include "internal.h"

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Device bytes a context for 2^n variables is expected to hold after a proof,
// until one has been measured: the polynomial (when uploaded), the rate-1/2
// codeword, its four-step buffer and Merkle tree, the folded sumcheck arrays
// and the eq tables; about 40 bytes per coefficient plus 16 MiB.
size_t estimate_bytes(uint32_t n, int on_device) {
  size_t coeffs = (size_t)1 << n;
  return coeffs * (on_device ? 36 : 40) + ((size_t)16 << 20);
}

uint32_t size_key(const whirgpu_job* j) { return 2 * j->num_variables + (j->evals_on_device ? 1u : 0u); }

// Jobs that can share a group: same size, constraint count and input kind.
bool same_shape(const whirgpu_job* a, const whirgpu_job* b) {
  return a->num_variables == b->num_variables && a->npoints == b->npoints &&
         (a->evals_on_device != 0) == (b->evals_on_device != 0);
}

constexpr int GROUP_LIMIT = 16;   // the library's group size limit

struct Pending {
  whirgpu_job* job;
  double submitted;
};

// One worker thread's contexts.  `ctxs` is touched by its own thread only;
// `bytes` (device bytes of each context) is read and written under the
// server's lock.
struct Worker {
  std::map<uint32_t, whirgpu_ctx*> ctxs;
  std::map<uint32_t, size_t> bytes;
  bool idle = false;   // waiting with nothing it may take (under the server's lock)
};

}  // namespace

struct whirgpu_server {
  int device = 0;
  whirgpu_params params{};
  whirgpu_poseidon2 perms{};
  size_t budget = 0;
  size_t held = 0;        // device bytes of every worker's contexts
  size_t reserved = 0;    // growth the running jobs may still add
  size_t releasing = 0;   // bytes of contexts being destroyed (still in `held`)
  size_t running = 0;      // jobs running
  int group_max = GROUP_LIMIT;
  bool self_check = false;   // verify every proof on the host before publishing it
  bool stop = false;
  std::map<uint32_t, size_t> peak;   // largest context bytes per proof measured per job size (size_key)
  std::map<uint32_t, whirgpu_tuning> tuning;   // launch shapes per number of variables
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Pending> queue;
  std::vector<std::unique_ptr<Worker>> workers;
  std::vector<std::thread> threads;
};

namespace {

// Bytes a group of k jobs like j is expected to add to worker w's contexts
// (under the lock).
size_t growth(const whirgpu_server* s, const Worker& w, const whirgpu_job* j, size_t k) {
  auto pk = s->peak.find(size_key(j));
  const size_t per = pk != s->peak.end() ? pk->second : estimate_bytes(j->num_variables, j->evals_on_device);
  const size_t expect = per * k;
  auto it = w.bytes.find(j->num_variables);
  const size_t have = it != w.bytes.end() ? it->second : 0;
  return expect > have ? expect - have : 0;
}

void set_error(whirgpu_job* j) {
  strncpy(j->error, whirgpu_last_error(), sizeof j->error - 1);
  j->error[sizeof j->error - 1] = 0;
}

// Proves the jobs (one shape) as one group on the worker's context for their
// size: statements created and their constraints evaluated for all jobs at
// once, then whirgpu_prove_group.  Returns each job's status (a failed job's
// error text is in its record); the caller publishes them.
std::vector<int> run_group(whirgpu_server* s, Worker* w, const std::vector<whirgpu_job*>& js, whirgpu_ctx** used) {
  const uint32_t B = (uint32_t)js.size();
  const uint32_t n = js[0]->num_variables;
  whirgpu_ctx* ctx = nullptr;
  auto it = w->ctxs.find(n);
  int rc = WHIRGPU_OK;
  if (it == w->ctxs.end()) {
    whirgpu_params p = s->params;
    p.num_variables = n;
    rc = whirgpu_create(s->device, &p, &s->perms, &ctx);
    if (rc == WHIRGPU_OK) {
      w->ctxs[n] = ctx;
      bool tuned = false;
      whirgpu_tuning t{};
      {
        std::lock_guard<std::mutex> lk(s->mu);
        auto ti = s->tuning.find(n);
        if (ti != s->tuning.end()) {
          t = ti->second;
          tuned = true;
        }
      }
      if (tuned) (void)whirgpu_set_tuning(ctx, &t);   // refused: the built-in shapes stay
    }
  } else {
    ctx = it->second;
  }
  *used = ctx;
  std::vector<whirgpu_statement*> sts(B, nullptr);
  std::vector<whirgpu_proof*> proofs(B, nullptr);
  std::vector<whirgpu_challenger> chs(B);
  for (uint32_t b = 0; rc == WHIRGPU_OK && b < B; b++)
    rc = whirgpu_statement_create(ctx, js[b]->evals, js[b]->evals_on_device, &sts[b]);
  std::vector<uint32_t> pts, evs(4 * B);
  std::vector<std::vector<uint32_t>> job_evals(B, std::vector<uint32_t>(4 * (size_t)js[0]->npoints + 4));
  for (uint32_t k = 0; rc == WHIRGPU_OK && k < js[0]->npoints; k++) {
    pts.assign((size_t)4 * n * B, 0);
    for (uint32_t b = 0; b < B; b++)
      memcpy(pts.data() + (size_t)4 * n * b, js[b]->points + (size_t)4 * n * k, (size_t)16 * n);
    rc = whirgpu_statement_evaluate_group(B, sts.data(), pts.data(), evs.data());
    for (uint32_t b = 0; rc == WHIRGPU_OK && b < B; b++) {
      memcpy(job_evals[b].data() + 4 * k, evs.data() + 4 * b, 16);
      if (js[b]->evals_out) memcpy(js[b]->evals_out + 4 * k, evs.data() + 4 * b, 16);
    }
  }
  for (uint32_t b = 0; rc == WHIRGPU_OK && b < B; b++) {
    rc = whirgpu_challenger_new(&chs[b]);
    if (rc == WHIRGPU_OK) rc = whirgpu_challenger_observe_domain_separator(ctx, &chs[b]);
    if (rc == WHIRGPU_OK) rc = whirgpu_proof_create(ctx, &proofs[b]);
  }
  if (rc == WHIRGPU_OK) rc = whirgpu_prove_group(ctx, B, sts.data(), proofs.data(), chs.data());
  bool check;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    check = s->self_check;
  }
  std::vector<int> status(B);
  for (uint32_t b = 0; b < B; b++) {
    whirgpu_job* j = js[b];
    int r = rc;
    if (r == WHIRGPU_OK && check) {
      // the verifier's own transcript from the domain separator, and the
      // evaluations the statement produced
      whirgpu_challenger vch;
      r = whirgpu_challenger_new(&vch);
      if (r == WHIRGPU_OK) r = whirgpu_challenger_observe_domain_separator(ctx, &vch);
      if (r == WHIRGPU_OK) r = whirgpu_verify(ctx, proofs[b], &vch, j->npoints, j->points, job_evals[b].data());
    }
    if (r == WHIRGPU_OK) r = whirgpu_proof_serialize(proofs[b], j->proof_out, j->proof_cap, &j->proof_len);
    if (r != WHIRGPU_OK) set_error(j);
    else j->error[0] = 0;
    status[b] = r;
    if (proofs[b]) whirgpu_proof_destroy(proofs[b]);
    if (sts[b]) whirgpu_statement_destroy(sts[b]);
  }
  return status;
}

// Bytes held by contexts (of any worker) for sizes other than n (under the lock).
size_t other_size_bytes(const whirgpu_server* s, uint32_t n) {
  size_t b = 0;
  for (const auto& w : s->workers)
    for (const auto& kv : w->bytes)
      if (kv.first != n) b += kv.second;
  return b;
}

// Whether any worker holds a context for size n (under the lock).
bool size_held(const whirgpu_server* s, uint32_t n) {
  for (const auto& w : s->workers)
    if (w->bytes.count(n)) return true;
  return false;
}

// Whether a worker other than w that is idle holds a context for size n.
bool held_by_idle(const whirgpu_server* s, const Worker& w, uint32_t n) {
  for (const auto& o : s->workers)
    if (o.get() != &w && o->idle && o->bytes.count(n)) return true;
  return false;
}

// Size affinity (under the lock).  Among the jobs submitted within
// AFFINITY_MS of the head (jobs submitted together, such as a batch), a
// worker takes the oldest of a size it holds a context for, so its context
// and pooled buffers are reused; else the oldest whose size no idle worker
// holds (an idle holder takes it at once, without creating a context or
// allocating); else nothing (end(): the worker waits and the holders take the
// jobs).  Only idle workers are deferred to, and an idle worker with a job of
// its own size in the window never waits, so a job is always taken.
constexpr double AFFINITY_MS = 5.0;
std::deque<Pending>::iterator pick_job(whirgpu_server* s, const Worker& w) {
  const double horizon = s->queue.front().submitted + AFFINITY_MS;
  for (auto it = s->queue.begin(); it != s->queue.end() && it->submitted <= horizon; ++it)
    if (w.bytes.count(it->job->num_variables)) return it;
  for (auto it = s->queue.begin(); it != s->queue.end() && it->submitted <= horizon; ++it)
    if (!held_by_idle(s, w, it->job->num_variables)) return it;
  return s->queue.end();
}

// Reserved growth plus a job's own growth fits in the device memory free now.
bool fits_free_memory(const whirgpu_server* s, size_t g) {
  if (g == 0) return true;
  size_t free_b = 0, total_b = 0;
  if (whirgpu_internal_mem_info(s->device, &free_b, &total_b) != WHIRGPU_OK) return false;
  return s->reserved + g <= free_b;
}

void worker(whirgpu_server* s, Worker* w) {
  for (;;) {
    std::vector<Pending> grp;
    size_t g = 0;
    bool more = false;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      bool take = false, alone = false;
      std::deque<Pending>::iterator pick;
      while (!take) {
        if (s->queue.empty()) {
          if (s->stop) break;
          w->idle = true;
          s->cv_work.wait(lk);
          w->idle = false;
          continue;
        }
        pick = pick_job(s, *w);
        if (pick == s->queue.end()) {   // left to idle workers holding those sizes
          w->idle = true;
          s->cv_work.wait(lk);
          w->idle = false;
          continue;
        }
        const whirgpu_job* j = pick->job;
        g = growth(s, *w, j, 1);
        if (s->held + s->reserved + g <= s->budget && fits_free_memory(s, g)) {
          take = true;
          break;
        }
        // Does not fit: release this worker's contexts of other sizes, then look again.
        // Their bytes stay in `held` (and count as `releasing`) until the
        // device memory is freed, outside the lock.
        std::vector<whirgpu_ctx*> drop;
        size_t dropped = 0;
        for (auto it = w->ctxs.begin(); it != w->ctxs.end();) {
          if (it->first != j->num_variables) {
            drop.push_back(it->second);
            dropped += w->bytes[it->first];
            w->bytes.erase(it->first);
            it = w->ctxs.erase(it);
          } else {
            ++it;
          }
        }
        if (drop.empty()) {
          // Nothing running and nothing held for other sizes or being
          // released: the job runs alone even if it exceeds the budget, on
          // this worker if it holds a context of the job's size or no worker
          // does (otherwise the worker that holds one takes it).  Any other
          // case waits for a job to finish or a worker to release memory.
          const uint32_t n = j->num_variables;
          if (s->running == 0 && s->releasing == 0 && other_size_bytes(s, n) == 0 &&
              (w->bytes.count(n) || !size_held(s, n))) {
            take = alone = true;
            break;
          }
          s->cv_work.wait(lk);
          continue;
        }
        s->releasing += dropped;
        lk.unlock();
        for (whirgpu_ctx* c : drop) whirgpu_destroy(c);
        lk.lock();
        s->releasing -= dropped;
        s->held -= dropped;
        s->cv_work.notify_all();   // less is held: others may fit now
      }
      if (!take) break;   // stopped with an empty queue
      grp.push_back(*pick);
      s->queue.erase(pick);
      // queued jobs of the same shape join the group while it fits the budget
      for (auto it = s->queue.begin(); !alone && it != s->queue.end() && (int)grp.size() < s->group_max;) {
        if (!same_shape(it->job, grp[0].job)) {
          ++it;
          continue;
        }
        const size_t g2 = growth(s, *w, grp[0].job, grp.size() + 1);
        if (s->held + s->reserved + g2 > s->budget || !fits_free_memory(s, g2)) break;
        g = g2;
        grp.push_back(*it);
        it = s->queue.erase(it);
      }
      s->reserved += g;
      s->running += grp.size();
      more = !s->queue.empty();
    }
    if (more) s->cv_work.notify_all();   // jobs left behind may wait for this worker to be busy
    const double t0 = now_ms();
    std::vector<whirgpu_job*> js;
    for (auto& p : grp) {
      p.job->queue_ms = t0 - p.submitted;
      js.push_back(p.job);
    }
    whirgpu_ctx* used = nullptr;
    std::vector<int> status = run_group(s, w, js, &used);
    if (js.size() > 1) {
      // a failed group is retried job by job, so one bad job fails alone
      bool failed = false;
      for (int r : status) failed = failed || r != WHIRGPU_OK;
      if (failed)
        for (size_t b = 0; b < js.size(); b++) status[b] = run_group(s, w, {js[b]}, &used)[0];
    }
    const size_t now_bytes = used ? whirgpu_internal_ctx_bytes(used) : 0;
    const double t1 = now_ms();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->reserved -= g;
      s->running -= grp.size();
      if (used) {
        size_t& b = w->bytes[js[0]->num_variables];
        s->held = s->held - b + now_bytes;
        b = now_bytes;
        size_t& pk = s->peak[size_key(js[0])];
        const size_t per = now_bytes / js.size();
        if (per > pk) pk = per;
      }
      for (size_t b = 0; b < js.size(); b++) {
        grp[b].job->latency_ms = t1 - grp[b].submitted;
        grp[b].job->status = status[b];
      }
    }
    s->cv_done.notify_all();
    s->cv_work.notify_all();
  }
  size_t dropped = 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto& kv : w->bytes) dropped += kv.second;
    w->bytes.clear();
  }
  for (auto& kv : w->ctxs) whirgpu_destroy(kv.second);
  w->ctxs.clear();
  std::lock_guard<std::mutex> lk(s->mu);
  s->held -= dropped;
}

}  // namespace

extern "C" {

int whirgpu_server_create(int device, const whirgpu_params* params, const whirgpu_poseidon2* perms, int workers,
                          size_t hbm_budget, whirgpu_server** out) {
  if (!params || !perms || !out || workers < 1 || workers > 256) return WHIRGPU_BAD_PARAMS;
  // jobs are proved in groups, whose transcripts run on the device: the
  // duplex challenger's, so the Keccak MMCS is proved one call at a time
  // (whirgpu_commit_keccak / whirgpu_prove_keccak) and not served here
  if (params->mmcs != WHIRGPU_MMCS_POSEIDON2) return WHIRGPU_BAD_PARAMS;
  size_t free_b = 0, total_b = 0;
  int rc = whirgpu_internal_mem_info(device, &free_b, &total_b);
  if (rc) return rc;
  auto* s = new whirgpu_server();
  s->device = device;
  s->params = *params;
  s->perms = *perms;
  s->budget = hbm_budget ? hbm_budget : free_b / 10 * 8;
  for (int i = 0; i < workers; i++) s->workers.emplace_back(new Worker());
  for (int i = 0; i < workers; i++) s->threads.emplace_back(worker, s, s->workers[i].get());
  *out = s;
  return WHIRGPU_OK;
}

int whirgpu_server_submit(whirgpu_server* s, whirgpu_job* job) {
  if (!s || !job || !job->evals || (job->npoints && !job->points) || !job->proof_out) return WHIRGPU_BAD_PARAMS;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return WHIRGPU_BAD_STATE;
    job->status = WHIRGPU_PENDING;
    job->proof_len = 0;
    job->error[0] = 0;
    s->queue.push_back({job, now_ms()});
  }
  s->cv_work.notify_all();
  return WHIRGPU_OK;
}

int whirgpu_server_submit_many(whirgpu_server* s, whirgpu_job* const* jobs, size_t count) {
  if (!s || (!jobs && count)) return WHIRGPU_BAD_PARAMS;
  for (size_t i = 0; i < count; i++) {
    const whirgpu_job* j = jobs[i];
    if (!j || !j->evals || (j->npoints && !j->points) || !j->proof_out) return WHIRGPU_BAD_PARAMS;
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return WHIRGPU_BAD_STATE;
    const double t = now_ms();
    // jobs submitted together are queued largest first (the longest-processing-
    // time rule: the small ones fill in at the end), in submission order within a size
    std::vector<whirgpu_job*> order(jobs, jobs + count);
    std::stable_sort(order.begin(), order.end(), [](const whirgpu_job* a, const whirgpu_job* b) {
      return a->num_variables > b->num_variables;
    });
    for (whirgpu_job* j : order) {
      j->status = WHIRGPU_PENDING;
      j->proof_len = 0;
      j->error[0] = 0;
      s->queue.push_back({j, t});
    }
  }
  s->cv_work.notify_all();
  return WHIRGPU_OK;
}

int whirgpu_server_set_group_max(whirgpu_server* s, int group_max) {
  if (!s || group_max < 1 || group_max > GROUP_LIMIT) return WHIRGPU_BAD_PARAMS;
  std::lock_guard<std::mutex> lk(s->mu);
  s->group_max = group_max;
  return WHIRGPU_OK;
}

int whirgpu_server_set_tuning(whirgpu_server* s, uint32_t num_variables, const whirgpu_tuning* t) {
  if (!s || !t) return WHIRGPU_BAD_PARAMS;
  std::lock_guard<std::mutex> lk(s->mu);
  s->tuning[num_variables] = *t;
  return WHIRGPU_OK;
}

int whirgpu_server_set_self_check(whirgpu_server* s, int on) {
  if (!s) return WHIRGPU_BAD_PARAMS;
  std::lock_guard<std::mutex> lk(s->mu);
  s->self_check = on != 0;
  return WHIRGPU_OK;
}

int whirgpu_server_wait(whirgpu_server* s, whirgpu_job* job, int timeout_ms) {
  if (!s || !job) return WHIRGPU_BAD_PARAMS;
  std::unique_lock<std::mutex> lk(s->mu);
  auto done = [&] { return job->status != WHIRGPU_PENDING; };
  if (timeout_ms < 0) s->cv_done.wait(lk, done);
  else s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done);
  return job->status;
}

int whirgpu_server_stats(whirgpu_server* s, size_t* queued, size_t* running, size_t* bytes_in_use) {
  if (!s) return WHIRGPU_BAD_PARAMS;
  std::lock_guard<std::mutex> lk(s->mu);
  if (queued) *queued = s->queue.size();
  if (running) *running = s->running;
  if (bytes_in_use) *bytes_in_use = s->held + s->reserved;
  return WHIRGPU_OK;
}

int whirgpu_server_destroy(whirgpu_server* s) {
  if (!s) return WHIRGPU_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& t : s->threads) t.join();
  delete s;
  return WHIRGPU_OK;
}

int whirgpu_prove_batch(int device, const whirgpu_params* params, const whirgpu_poseidon2* perms, whirgpu_job* jobs,
                        size_t njobs, int workers) {
  if (!jobs && njobs) return WHIRGPU_BAD_PARAMS;
  whirgpu_server* s = nullptr;
  int rc = whirgpu_server_create(device, params, perms, workers, 0, &s);
  if (rc) return rc;
  int first = WHIRGPU_OK;
  std::vector<whirgpu_job*> valid;
  for (size_t i = 0; i < njobs; i++) {
    const whirgpu_job& j = jobs[i];
    if (!j.evals || (j.npoints && !j.points) || !j.proof_out) {
      jobs[i].status = WHIRGPU_BAD_PARAMS;
      if (!first) first = WHIRGPU_BAD_PARAMS;
    } else {
      valid.push_back(&jobs[i]);
    }
  }
  int r = whirgpu_server_submit_many(s, valid.data(), valid.size());   // one queue update: groups can form
  if (r) {
    for (whirgpu_job* j : valid) j->status = r;
    if (!first) first = r;
  }
  for (size_t i = 0; i < njobs; i++) {
    int r = whirgpu_server_wait(s, &jobs[i], -1);   // returns at once for a job that failed to submit
    if (r && !first) first = r;
  }
  whirgpu_server_destroy(s);
  return first;
}

}  // extern "C"
