// Batched and served proving on one device: a FIFO of jobs and a pool of host
// worker threads.  Each worker keeps its own contexts (one per number of
// variables, each with its own HIP stream and buffer pool), so proofs from
// different workers overlap on the device: while one worker waits for a
// round's transcript values another worker's kernels run.
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
  size_t running = 0;
  bool stop = false;
  std::map<uint32_t, size_t> peak;   // largest context measured per job size (size_key)
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Pending> queue;
  std::vector<std::unique_ptr<Worker>> workers;
  std::vector<std::thread> threads;
};

namespace {

// Bytes job j is expected to add to worker w's contexts (under the lock).
size_t growth(const whirgpu_server* s, const Worker& w, const whirgpu_job* j) {
  auto pk = s->peak.find(size_key(j));
  const size_t expect = pk != s->peak.end() ? pk->second : estimate_bytes(j->num_variables, j->evals_on_device);
  auto it = w.bytes.find(j->num_variables);
  const size_t have = it != w.bytes.end() ? it->second : 0;
  return expect > have ? expect - have : 0;
}

int run_job(whirgpu_server* s, Worker* w, whirgpu_job* j, whirgpu_ctx** used) {
  whirgpu_ctx* ctx;
  auto it = w->ctxs.find(j->num_variables);
  if (it == w->ctxs.end()) {
    whirgpu_params p = s->params;
    p.num_variables = j->num_variables;
    int rc = whirgpu_create(s->device, &p, &s->perms, &ctx);
    if (rc) return rc;
    w->ctxs[j->num_variables] = ctx;
  } else {
    ctx = it->second;
  }
  *used = ctx;
  whirgpu_statement* st = nullptr;
  whirgpu_proof* proof = nullptr;
  whirgpu_prover_data* pd = nullptr;
  int rc = whirgpu_statement_create(ctx, j->evals, j->evals_on_device, &st);
  const uint32_t n = j->num_variables;
  for (uint32_t k = 0; rc == 0 && k < j->npoints; k++) {
    uint32_t ev[4];
    rc = whirgpu_statement_evaluate(st, j->points + (size_t)4 * n * k, ev);
    if (rc == 0 && j->evals_out) memcpy(j->evals_out + 4 * k, ev, sizeof ev);
  }
  whirgpu_challenger ch;
  if (rc == 0) rc = whirgpu_challenger_new(&ch);
  if (rc == 0) rc = whirgpu_challenger_observe_domain_separator(ctx, &ch);
  if (rc == 0) rc = whirgpu_proof_create(ctx, &proof);
  if (rc == 0) rc = whirgpu_commit(ctx, proof, &ch, st, &pd);
  if (rc == 0) {
    rc = whirgpu_prove(ctx, proof, &ch, st, pd);   // consumes pd
    pd = nullptr;
  }
  if (rc == 0) rc = whirgpu_proof_serialize(proof, j->proof_out, j->proof_cap, &j->proof_len);
  if (pd) whirgpu_prover_data_destroy(pd);
  if (proof) whirgpu_proof_destroy(proof);
  if (st) whirgpu_statement_destroy(st);
  return rc;
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

// Reserved growth plus a job's own growth fits in the device memory free now.
bool fits_free_memory(const whirgpu_server* s, size_t g) {
  if (g == 0) return true;
  size_t free_b = 0, total_b = 0;
  if (whirgpu_internal_mem_info(s->device, &free_b, &total_b) != WHIRGPU_OK) return false;
  return s->reserved + g <= free_b;
}

void worker(whirgpu_server* s, Worker* w) {
  for (;;) {
    Pending p{};
    size_t g = 0;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      bool take = false;
      while (!take) {
        if (s->queue.empty()) {
          if (s->stop) break;
          s->cv_work.wait(lk);
          continue;
        }
        const whirgpu_job* j = s->queue.front().job;
        g = growth(s, *w, j);
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
            take = true;
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
      p = s->queue.front();
      s->queue.pop_front();
      s->reserved += g;
      s->running++;
    }
    double t0 = now_ms();
    p.job->queue_ms = t0 - p.submitted;
    whirgpu_ctx* used = nullptr;
    int rc = run_job(s, w, p.job, &used);
    if (rc) {
      strncpy(p.job->error, whirgpu_last_error(), sizeof p.job->error - 1);
      p.job->error[sizeof p.job->error - 1] = 0;
    } else {
      p.job->error[0] = 0;
    }
    const size_t now_bytes = used ? whirgpu_internal_ctx_bytes(used) : 0;
    double t1 = now_ms();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->reserved -= g;
      s->running--;
      if (used) {
        size_t& b = w->bytes[p.job->num_variables];
        s->held = s->held - b + now_bytes;
        b = now_bytes;
        size_t& pk = s->peak[size_key(p.job)];
        if (now_bytes > pk) pk = now_bytes;
      }
      p.job->latency_ms = t1 - p.submitted;
      p.job->status = rc;
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
  for (size_t i = 0; i < njobs; i++) {
    int r = whirgpu_server_submit(s, &jobs[i]);
    if (r) {
      jobs[i].status = r;
      if (!first) first = r;
    }
  }
  for (size_t i = 0; i < njobs; i++) {
    int r = whirgpu_server_wait(s, &jobs[i], -1);   // returns at once for a job that failed to submit
    if (r && !first) first = r;
  }
  whirgpu_server_destroy(s);
  return first;
}

}  // extern "C"
