// Serving runtime for encrypted queries against device-resident databases
// (no counterpart in the reference, which has no server).
//
// A worker thread runs a step loop with continuous batching:
//   1. admit queued requests while a slot is free and the per-query HBM
//      footprint (expanded ciphertexts + answer) fits the budget; the query
//      ciphertext is copied into the slot and its expansion starts at
//      level 0;
//   2. advance every active request by one stage: all requests at
//      expansion level j share one shg_oblivious_expand_level call, all
//      fully expanded requests of one database share one
//      shg_pir_inner_product call;
//   3. wait for the stream, copy finished answers to host memory, release
//      their slots and wake their waiters.
// Requests that arrive while others are on the GPU join at the next step,
// so their work overlaps on the device with the requests already running.
// Keys and databases stay resident; every result depends only on its own
// query, so it is the same alone and in any mix.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "shell_gpu.h"
#include "shg_kernels.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  long long id = 0;
  int db = 0;
  std::vector<uint64_t> query;   // 2 * L * n words (host copy)
  std::vector<uint64_t> answer;  // C * 2 * L * n words when done
  Clock::time_point submitted, finished;
  bool done = false;
  int rc = SHG_OK;
  std::string error;  // the worker's shg_last_error() message when rc != SHG_OK
};

struct Slot {
  uint64_t* expanded = nullptr;  // 2^levels ciphertexts
  uint64_t* answer = nullptr;    // max_columns ciphertexts
  std::shared_ptr<Request> req;
  int stage = 0;                 // expansion level, or levels = inner product
};

struct Db {
  const uint64_t* exported;
  int columns;
};

}  // namespace

struct shg_pir_server {
  shg_context* ctx = nullptr;
  std::vector<int> primes;
  std::vector<int> log_bs;
  std::vector<const uint64_t*> key_as, key_bs;
  std::vector<shg_ks_key*> keys;  // the level keys, prepared at creation
  int levels = 0;
  int log_n = 0;
  size_t ct_words = 0;  // 2 * L * n
  shg_pir_server_config cfg{};
  std::vector<Db> dbs;
  int max_columns = 1;
  std::vector<Slot> slots;
  size_t footprint = 0;  // bytes per admitted request
  int capacity = 0;      // slots allowed by max_active and the HBM budget

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> pending;
  std::map<long long, std::shared_ptr<Request>> requests;
  long long next_id = 1;
  bool stop = false;
  bool started = false;
  std::thread worker;
  long long steps = 0, stage_runs = 0, batched_items = 0;

  void run();
  int step(std::vector<Slot*>& active);
};

namespace {

int slot_alloc(shg_pir_server* s, Slot& sl) {
  if (sl.expanded) return SHG_OK;
  const size_t exp_bytes = ((size_t)1 << s->levels) * s->ct_words * 8;
  const size_t ans_bytes = (size_t)s->max_columns * s->ct_words * 8;
  int rc = shg_alloc(s->ctx, exp_bytes, (void**)&sl.expanded);
  if (rc) return rc;
  rc = shg_alloc(s->ctx, ans_bytes, (void**)&sl.answer);
  if (rc) {
    shg_free(s->ctx, sl.expanded);
    sl.expanded = nullptr;
  }
  return rc;
}

}  // namespace

int shg_pir_server::step(std::vector<Slot*>& active) {
  const int L = (int)primes.size();
  // Expansion: one call per level present.
  for (int j = 0; j < levels; ++j) {
    std::vector<uint64_t*> ptrs;
    for (Slot* sl : active)
      if (sl->stage == j) ptrs.push_back(sl->expanded);
    if (ptrs.empty()) continue;
    int rc = shg_oblivious_expand_level_key(ctx, ptrs.data(), (int)ptrs.size(), j, keys[j],
                                            log_bs.data(), L, primes.data());
    if (rc) return rc;
    ++stage_runs;
    batched_items += (long long)ptrs.size();
  }
  // Inner products: one call per database present.
  for (size_t d = 0; d < dbs.size(); ++d) {
    std::vector<const uint64_t*> ex;
    std::vector<uint64_t*> out;
    for (Slot* sl : active)
      if (sl->stage == levels && sl->req->db == (int)d) {
        ex.push_back(sl->expanded);
        out.push_back(sl->answer);
      }
    if (ex.empty()) continue;
    int rc = shg_pir_inner_product(ctx, ex.data(), (int)ex.size(), levels, dbs[d].exported,
                                   dbs[d].columns, out.data(), L, primes.data());
    if (rc) return rc;
    ++stage_runs;
    batched_items += (long long)ex.size();
  }
  return shg_sync(ctx);
}

void shg_pir_server::run() {
  std::vector<Slot*> active;
  for (;;) {
    // 1. Admission.
    {
      std::unique_lock<std::mutex> lock(mu);
      if (active.empty())
        cv_work.wait(lock, [&] { return stop || !pending.empty(); });
      if (stop) break;
      // Cohorts: while requests are on the device, newcomers wait until
      // `cohort` of them can start together, so that they advance through
      // the stages (and share the batched calls) as one group.  An idle
      // device admits whatever is queued.
      const int cohort = std::max(1, std::min(cfg.cohort, capacity));
      const int free_slots = capacity - (int)active.size();
      const bool admit = active.empty() ||
                         (free_slots >= cohort && (int)pending.size() >= cohort);
      while (admit && !pending.empty() && (int)active.size() < capacity) {
        Slot* free_slot = nullptr;
        for (auto& sl : slots)
          if (!sl.req) {
            free_slot = &sl;
            break;
          }
        if (!free_slot) break;
        auto req = pending.front();
        pending.pop_front();
        int rc = slot_alloc(this, *free_slot);
        if (rc == SHG_OK)
          rc = shg_upload(ctx, free_slot->expanded, req->query.data(), ct_words * 8);
        if (rc) {
          req->rc = rc;
          req->error = shg_last_error();
          req->done = true;
          req->finished = Clock::now();
          cv_done.notify_all();
          continue;
        }
        free_slot->req = req;
        free_slot->stage = 0;
        active.push_back(free_slot);
      }
    }
    if (active.empty()) continue;
    // 2. One stage for every active request.
    const int rc = step(active);
    ++steps;
    // 3. Completion.
    std::vector<Slot*> still;
    for (Slot* sl : active) {
      const bool finished = rc != SHG_OK || sl->stage == levels;
      if (!finished) {
        ++sl->stage;
        still.push_back(sl);
        continue;
      }
      auto& req = sl->req;
      int r = rc;
      if (r == SHG_OK) {
        const int C = dbs[req->db].columns;
        req->answer.resize((size_t)C * ct_words);
        r = shg_download(ctx, req->answer.data(), sl->answer, req->answer.size() * 8);
      }
      const std::string msg = r != SHG_OK ? std::string(shg_last_error()) : std::string();
      std::lock_guard<std::mutex> lock(mu);
      req->rc = r;
      req->error = msg;
      req->done = true;
      req->finished = Clock::now();
      sl->req.reset();
      cv_done.notify_all();
    }
    active.swap(still);
  }
  // Stopped with requests still on the device: fail them.
  std::lock_guard<std::mutex> lock(mu);
  for (Slot* sl : active) {
    sl->req->rc = SHG_E_FAILED_PRECONDITION;
    sl->req->error = "the server stopped before the request finished.";
    sl->req->done = true;
    sl->req.reset();
  }
  cv_done.notify_all();
}

extern "C" {

int shg_pir_server_create(shg_context* ctx, const int* primes, int num_limbs,
                          const int* host_log_bs, const uint64_t* const* key_as,
                          const uint64_t* const* key_bs, const shg_pir_server_config* cfg,
                          shg_pir_server** out) {
  using shg::set_error;
  if (!ctx || !primes || !host_log_bs || !cfg || !out || num_limbs <= 0)
    return set_error(SHG_E_INVALID_ARGUMENT, "null argument or no moduli.");
  if (cfg->log_n < 4 || cfg->log_n > 14)
    return set_error(SHG_E_INVALID_ARGUMENT, "log_n must be in [4, 14].");
  if (cfg->levels < 1 || cfg->levels > cfg->log_n)
    return set_error(SHG_E_INVALID_ARGUMENT, "levels must be in [1, log_n].");
  if (cfg->max_active < 1) return set_error(SHG_E_INVALID_ARGUMENT, "max_active must be positive.");
  if (cfg->cohort < 0) return set_error(SHG_E_INVALID_ARGUMENT, "cohort must be non-negative.");
  if (!key_as || !key_bs) return set_error(SHG_E_INVALID_ARGUMENT, "one Galois key per level is required.");
  auto s = std::make_unique<shg_pir_server>();
  s->ctx = ctx;
  s->primes.assign(primes, primes + num_limbs);
  s->log_bs.assign(host_log_bs, host_log_bs + num_limbs);
  s->key_as.assign(key_as, key_as + cfg->levels);
  s->key_bs.assign(key_bs, key_bs + cfg->levels);
  // Each level's key is prepared once (FP64-form copies for the digit-plane
  // key switch); its dimension is the gadget's, sum_i ceil(log q_i / log b_i).
  int k = 0;
  for (int i = 0; i < num_limbs; ++i) {
    uint64_t params[10];
    int rc = shg_prime_info(ctx, primes[i], params, nullptr);
    if (rc) return rc;
    if (host_log_bs[i] <= 0) return set_error(SHG_E_INVALID_ARGUMENT, "gadget base log must be positive.");
    k += (int)((params[1] + host_log_bs[i] - 1) / host_log_bs[i]);
  }
  for (int j = 0; j < cfg->levels; ++j) {
    shg_ks_key* key = nullptr;
    int rc = shg_ks_key_create(ctx, key_as[j], key_bs[j], k, num_limbs, primes, &key);
    if (rc) {
      for (shg_ks_key* made : s->keys) (void)shg_ks_key_destroy(ctx, made);
      return rc;
    }
    s->keys.push_back(key);
  }
  s->levels = cfg->levels;
  s->log_n = cfg->log_n;
  s->cfg = *cfg;
  s->ct_words = (size_t)2 * num_limbs << cfg->log_n;
  s->slots.resize(cfg->max_active);
  *out = s.release();
  return SHG_OK;
}

int shg_pir_server_add_db(shg_pir_server* s, const uint64_t* db_exported, int num_columns,
                          int* db_id) {
  if (!s || !db_exported || num_columns <= 0 || !db_id)
    return shg::set_error(SHG_E_INVALID_ARGUMENT, "invalid database.");
  std::lock_guard<std::mutex> lock(s->mu);
  if (s->started)
    return shg::set_error(SHG_E_FAILED_PRECONDITION, "databases must be added before start.");
  s->dbs.push_back({db_exported, num_columns});
  s->max_columns = num_columns > s->max_columns ? num_columns : s->max_columns;
  *db_id = (int)s->dbs.size() - 1;
  return SHG_OK;
}

int shg_pir_server_start(shg_pir_server* s) {
  if (!s) return SHG_E_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(s->mu);
  if (s->started) return SHG_OK;
  if (s->dbs.empty()) return shg::set_error(SHG_E_FAILED_PRECONDITION, "no database.");
  s->footprint = (((size_t)1 << s->levels) + (size_t)s->max_columns) * s->ct_words * 8;
  const size_t by_budget =
      s->cfg.hbm_budget_bytes ? s->cfg.hbm_budget_bytes / s->footprint : (size_t)s->cfg.max_active;
  s->capacity = (int)(by_budget < (size_t)s->cfg.max_active ? by_budget : s->cfg.max_active);
  if (s->capacity < 1)
    return shg::set_error(SHG_E_NO_MEMORY, "the HBM budget does not fit one request.");
  s->started = true;
  s->worker = std::thread([s] { s->run(); });
  return SHG_OK;
}

int shg_pir_server_submit(shg_pir_server* s, int db_id, const uint64_t* host_query,
                          long long* request) {
  if (!s || !host_query || !request) return SHG_E_INVALID_ARGUMENT;
  auto req = std::make_shared<Request>();
  req->db = db_id;
  req->query.assign(host_query, host_query + s->ct_words);
  std::lock_guard<std::mutex> lock(s->mu);
  if (!s->started || s->stop)
    return shg::set_error(SHG_E_FAILED_PRECONDITION, "the server is not running.");
  if (db_id < 0 || db_id >= (int)s->dbs.size())
    return shg::set_error(SHG_E_INVALID_ARGUMENT, "unknown database.");
  req->id = s->next_id++;
  req->submitted = Clock::now();
  s->requests[req->id] = req;
  s->pending.push_back(req);
  *request = req->id;
  s->cv_work.notify_one();
  return SHG_OK;
}

int shg_pir_server_wait(shg_pir_server* s, long long request, uint64_t* host_answer,
                        double* latency_ms) {
  if (!s) return SHG_E_INVALID_ARGUMENT;
  std::shared_ptr<Request> req;
  {
    std::unique_lock<std::mutex> lock(s->mu);
    auto it = s->requests.find(request);
    if (it == s->requests.end()) return shg::set_error(SHG_E_INVALID_ARGUMENT, "unknown request.");
    req = it->second;
    s->cv_done.wait(lock, [&] { return req->done; });
    s->requests.erase(it);
  }
  // The worker thread recorded its error message with the request; report it
  // on the waiting thread, whose shg_last_error() the caller reads.
  if (req->rc) return shg::set_error(req->rc, req->error.c_str());
  if (host_answer) std::memcpy(host_answer, req->answer.data(), req->answer.size() * 8);
  if (latency_ms)
    *latency_ms = std::chrono::duration<double, std::milli>(req->finished - req->submitted).count();
  return SHG_OK;
}

int shg_pir_server_stats(shg_pir_server* s, long long* steps, long long* stage_runs,
                         long long* batched_items, int* capacity) {
  if (!s) return SHG_E_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(s->mu);
  if (steps) *steps = s->steps;
  if (stage_runs) *stage_runs = s->stage_runs;
  if (batched_items) *batched_items = s->batched_items;
  if (capacity) *capacity = s->capacity;
  return SHG_OK;
}

int shg_pir_server_load_test(shg_pir_server* s, int db_id, const uint64_t* host_queries,
                             int num_distinct, int num_requests, double rate_per_s,
                             uint64_t seed, double* latencies_ms, double* wall_s) {
  if (!s || !host_queries || num_distinct <= 0 || num_requests <= 0 || !latencies_ms ||
      rate_per_s <= 0)
    return SHG_E_INVALID_ARGUMENT;
  std::vector<long long> ids(num_requests, 0);
  int submit_rc = SHG_OK;
  const auto t0 = Clock::now();
  // Poisson arrivals: exponential gaps at rate_per_s, from a thread of
  // their own so that submission never waits for the GPU.
  std::thread gen([&] {
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(rate_per_s);
    auto next = Clock::now();
    for (int i = 0; i < num_requests; ++i) {
      next += std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(gap(rng)));
      std::this_thread::sleep_until(next);
      const uint64_t* q = host_queries + (size_t)(i % num_distinct) * s->ct_words;
      int rc = shg_pir_server_submit(s, db_id, q, &ids[i]);
      if (rc) {
        submit_rc = rc;
        return;
      }
    }
  });
  gen.join();
  if (submit_rc) return submit_rc;
  int rc = SHG_OK;
  for (int i = 0; i < num_requests; ++i) {
    int r = shg_pir_server_wait(s, ids[i], nullptr, &latencies_ms[i]);
    if (r && !rc) rc = r;
  }
  if (wall_s) *wall_s = std::chrono::duration<double>(Clock::now() - t0).count();
  return rc;
}

int shg_pir_server_destroy(shg_pir_server* s) {
  if (!s) return SHG_E_INVALID_ARGUMENT;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    s->stop = true;
    for (auto& r : s->pending) {
      r->rc = SHG_E_FAILED_PRECONDITION;
      r->error = "the server stopped before the request started.";
      r->done = true;
    }
    s->pending.clear();
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  s->cv_done.notify_all();
  for (auto& sl : s->slots) {
    if (sl.expanded) shg_free(s->ctx, sl.expanded);
    if (sl.answer) shg_free(s->ctx, sl.answer);
  }
  for (shg_ks_key* key : s->keys) (void)shg_ks_key_destroy(s->ctx, key);
  delete s;
  return SHG_OK;
}

}  // extern "C"
