// Batched and served proving on one device: a FIFO of jobs and a pool of host
// worker threads.  Each worker keeps its own contexts (one per number of
// variables, each with its own HIP stream and buffer pool), so proofs from
// different workers overlap on the device: while one worker waits for a
// round's two transcript values another worker's kernels run.  A job is
// admitted when a worker is free and its device-memory estimate fits in the
// budget beside the jobs already running (always admitted when nothing runs).
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/whir_gpu.h"

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Device bytes a proof of 2^n variables holds at its peak: the polynomial
// (when uploaded), the rate-1/2 codeword, its four-step buffer and Merkle tree,
// and the folded sumcheck arrays; about 40 bytes per coefficient.
size_t estimate_bytes(uint32_t n, int on_device) {
  size_t coeffs = (size_t)1 << n;
  return coeffs * (on_device ? 36 : 40) + ((size_t)16 << 20);
}

struct Pending {
  whirgpu_job* job;
  double submitted;
};

}  // namespace

struct whirgpu_server {
  int device = 0;
  whirgpu_params params{};
  whirgpu_poseidon2 perms{};
  size_t budget = 0;
  size_t in_use = 0;
  size_t running = 0;
  bool stop = false;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Pending> queue;
  std::vector<std::thread> threads;
};

namespace {

int run_job(whirgpu_server* s, std::map<uint32_t, whirgpu_ctx*>& ctxs, whirgpu_job* j) {
  whirgpu_ctx* ctx;
  auto it = ctxs.find(j->num_variables);
  if (it == ctxs.end()) {
    whirgpu_params p = s->params;
    p.num_variables = j->num_variables;
    int rc = whirgpu_create(s->device, &p, &s->perms, &ctx);
    if (rc) return rc;
    ctxs[j->num_variables] = ctx;
  } else {
    ctx = it->second;
  }
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

void worker(whirgpu_server* s) {
  std::map<uint32_t, whirgpu_ctx*> ctxs;
  for (;;) {
    Pending p;
    size_t est;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] {
        if (s->queue.empty()) return s->stop;
        const whirgpu_job* j = s->queue.front().job;
        size_t e = estimate_bytes(j->num_variables, j->evals_on_device);
        return s->running == 0 || s->in_use + e <= s->budget;
      });
      if (s->queue.empty()) break;
      p = s->queue.front();
      s->queue.pop_front();
      est = estimate_bytes(p.job->num_variables, p.job->evals_on_device);
      s->in_use += est;
      s->running++;
    }
    double t0 = now_ms();
    p.job->queue_ms = t0 - p.submitted;
    int rc = run_job(s, ctxs, p.job);
    double t1 = now_ms();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      p.job->latency_ms = t1 - p.submitted;
      p.job->status = rc;
      s->in_use -= est;
      s->running--;
    }
    s->cv_done.notify_all();
    s->cv_work.notify_all();
  }
  for (auto& kv : ctxs) whirgpu_destroy(kv.second);
}

}  // namespace

extern "C" {

int whirgpu_server_create(int device, const whirgpu_params* params, const whirgpu_poseidon2* perms, int workers,
                          size_t hbm_budget, whirgpu_server** out) {
  if (!params || !perms || !out || workers < 1 || workers > 256) return WHIRGPU_BAD_PARAMS;
  if (hipSetDevice(device) != hipSuccess) return WHIRGPU_DEVICE;
  auto* s = new whirgpu_server();
  s->device = device;
  s->params = *params;
  s->perms = *perms;
  if (hbm_budget == 0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
      delete s;
      return WHIRGPU_DEVICE;
    }
    hbm_budget = free_b / 10 * 8;
  }
  s->budget = hbm_budget;
  for (int i = 0; i < workers; i++) s->threads.emplace_back(worker, s);
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
  if (bytes_in_use) *bytes_in_use = s->in_use;
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
