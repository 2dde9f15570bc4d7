// Serving loop: continuous batching of construction requests under an HBM
// budget, resident indices, and query batches answered across indices.
//
// Threads: one scheduler and `workers` worker threads, each worker with its
// own pwmhip_ctx (its own HIP stream and scratch arenas), so the batches of
// different workers overlap on the device.  All server state sits under one
// mutex; device work runs outside it.
//
// Budget: bytes_in_use = workers x scratch allowance (held while the server
// lives) + resident indices (levels + rank/select directories) + estimated
// output of builds in flight.  A build is admitted when its estimate fits.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "pwm_internal.hpp"

namespace {

using clk = std::chrono::steady_clock;

double ms_since(clk::time_point t0, clk::time_point t1) {
  return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

constexpr uint64_t kChunk = 4096;

// Device bytes of a structure's levels (as alloc_structure lays them out).
uint64_t output_bytes(uint64_t n, uint32_t L) { return n ? uint64_t(L) * pwm::level_stride(n) * 8 : 0; }

// Device bytes of the rank/select directories of a structure.
uint64_t rs_bytes(uint64_t n, uint32_t L) {
  auto al = [](uint64_t x) { return (x + 255) & ~uint64_t(255); };
  const uint64_t per = al(((n >> 16) + 2) * 8) + al(((n >> 9) + 1) * 2) + 2 * al((n / 8192 + 1) * 4);
  return L * per + al(L * 64) + al(L * 8) + 256;
}

// Scratch of one batched build of `chunks` chunks and `bytes` host text.
uint64_t scratch_bytes(uint64_t chunks, uint64_t texts, uint64_t text_bytes, uint32_t L) {
  const uint64_t nslot = L > 1 ? (2ull << (std::min<uint32_t>(L, 8) - 1)) - 2 : 1;
  const uint64_t temp = uint64_t(std::max<uint32_t>(L, 2) - 1) * (chunks * 64 + 64) * 8;
  const uint64_t tables = chunks * nslot * (4 + 8 * 4) + ((chunks + 1023) / 1024 + 1) * nslot * 8;
  return temp + tables + texts * 256 + text_bytes + (64ull << 20);
}

struct Block {  // device block shared by the structures of one batch
  pwmhip_ws owner{};
  uint64_t bytes = 0;
  int refs = 0;
};

struct Index {
  int status = 1;  // 1 pending, 0 ready, < 0 failed
  uint64_t n = 0, est = 0, bytes = 0;
  uint32_t width = 1, levels = 0, algo = 0, flags = 0;
  const void* text = nullptr;
  pwmhip_ws ws{};
  pwmhip_rs* rs = nullptr;
  std::shared_ptr<Block> block;
  clk::time_point t_submit;
  bool release_requested = false;
  uint64_t queries_pending = 0;
};

struct QueryReq {
  uint64_t id = 0, index = 0, n = 0;
  const uint32_t* op = nullptr;
  const uint64_t *a = nullptr, *b = nullptr;
  uint64_t* result = nullptr;
  clk::time_point t_submit;
};

struct Job {
  enum Kind { NONE, BUILD, QUERY } kind = NONE;
  std::vector<uint64_t> builds;   // index ids
  std::vector<QueryReq> queries;
};

}  // namespace

struct pwmhip_server {
  pwmhip_server_config cfg{};
  uint64_t allowance = 0;  // scratch held per worker
  std::mutex mu;
  std::condition_variable cv_sched, cv_done, cv_work;
  bool stop = false;     // destroy called: finish submitted work
  bool exiting = false;  // scheduler done: idle workers leave
  uint64_t next_id = 1;
  std::deque<uint64_t> pending_builds;
  std::deque<QueryReq> pending_queries;
  std::unordered_map<uint64_t, Index> indices;
  std::unordered_map<uint64_t, int> query_status;  // 1 pending, else final status
  uint64_t in_use = 0, peak = 0;

  struct Worker {
    pwmhip_ctx* ctx = nullptr;
    std::thread th;
    Job job;
    bool busy = false;
  };
  std::vector<std::unique_ptr<Worker>> workers;
  std::thread sched;

  // statistics
  std::vector<double> build_lat, query_lat;
  uint64_t builds_failed = 0, queries_done = 0, build_launches = 0, query_launches = 0, texts_launched = 0;

  void charge(int64_t delta) {
    in_use = uint64_t(int64_t(in_use) + delta);
    peak = std::max(peak, in_use);
  }
  void scheduler();
  void worker(Worker* w);
  void run_build(Worker* w, Job& job);
  void run_queries(Worker* w, Job& job);
  void fail_job(const Job& job, int st);  // under mu
  void drop_index(uint64_t id);  // under mu
};

void pwmhip_server::drop_index(uint64_t id) {
  auto it = indices.find(id);
  if (it == indices.end()) return;
  Index& x = it->second;
  pwmhip_ctx* c = workers[0]->ctx;
  if (x.rs) pwmhip_rs_free(c, x.rs);
  x.rs = nullptr;
  int64_t freed = int64_t(x.bytes);
  if (x.block && --x.block->refs == 0) {
    pwmhip_ws_free(c, &x.block->owner);
    freed += int64_t(x.block->bytes);
  }
  x.block.reset();
  // a structure outside a shared block owns its words (and host arrays)
  if (x.ws.alloc_bytes || x.ws.code_len || x.ws.code_word) pwmhip_ws_free(c, &x.ws);
  charge(-freed);
  indices.erase(it);
}

void pwmhip_server::scheduler() {
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    Worker* free_w = nullptr;
    for (auto& w : workers)
      if (!w->busy) {
        free_w = w.get();
        break;
      }
    bool dispatched = false;
    if (free_w) {
      // queries of ready indices first (they are short)
      Job qj;
      qj.kind = Job::QUERY;
      uint64_t nq = 0;
      for (auto it = pending_queries.begin(); it != pending_queries.end();) {
        auto ix = indices.find(it->index);
        const bool ready = ix != indices.end() && ix->second.status == 0;
        const bool failed = ix == indices.end() || ix->second.status < 0;
        if (failed) {
          query_status[it->id] = ix == indices.end() ? PWMHIP_ERR_INVALID : ix->second.status;
          if (ix != indices.end()) ix->second.queries_pending--;
          it = pending_queries.erase(it);
          cv_done.notify_all();
          continue;
        }
        if (ready && (nq == 0 || nq + it->n <= cfg.max_query_batch)) {
          nq += it->n;
          qj.queries.push_back(*it);
          it = pending_queries.erase(it);
          continue;
        }
        ++it;
      }
      if (!qj.queries.empty()) {
        free_w->job = std::move(qj);
        free_w->busy = true;
        dispatched = true;
      } else if (!pending_builds.empty()) {
        // the oldest waiting build fixes the shape; same-shape builds join it
        const Index& head = indices[pending_builds.front()];
        const uint32_t width = head.width, levels = head.levels, algo = head.algo, flags = head.flags;
        Job bj;
        bj.kind = Job::BUILD;
        uint64_t bytes = 0, chunks = 0;
        for (auto it = pending_builds.begin(); it != pending_builds.end();) {
          Index& x = indices[*it];
          const bool same = x.width == width && x.levels == levels && x.algo == algo && x.flags == flags;
          if (!same) {
            ++it;
            continue;
          }
          const uint64_t xb = x.n * x.width;
          const uint64_t xc = (x.n + kChunk - 1) / kChunk;
          if (!bj.builds.empty() &&
              (bj.builds.size() >= cfg.max_batch_texts || bytes + xb > cfg.max_batch_bytes ||
               scratch_bytes(chunks + xc, bj.builds.size() + 1, bytes + xb, levels) > allowance))
            break;
          if (in_use + x.est > cfg.hbm_budget_bytes) break;  // waits for memory
          charge(int64_t(x.est));
          bytes += xb;
          chunks += xc;
          bj.builds.push_back(*it);
          it = pending_builds.erase(it);
          if (levels > pwm::kMaxGroup) break;  // built one by one anyway
        }
        if (!bj.builds.empty()) {
          free_w->job = std::move(bj);
          free_w->busy = true;
          dispatched = true;
        }
      }
    }
    if (dispatched) {
      cv_work.notify_all();
      continue;
    }
    bool busy = false;
    for (auto& w : workers) busy |= w->busy;
    if (stop && !busy && pending_builds.empty() && pending_queries.empty()) break;
    if (stop && !busy && pending_queries.empty() && !pending_builds.empty()) {
      // shutting down and the budget holds the remaining builds back:
      // resident indices are no longer needed
      std::vector<uint64_t> ids;
      for (auto& kv : indices)
        if (kv.second.status != 1 && kv.second.queries_pending == 0) ids.push_back(kv.first);
      for (uint64_t id : ids) drop_index(id);
      if (!ids.empty()) continue;
    }
    cv_sched.wait(lk);
  }
  exiting = true;
  cv_work.notify_all();
}

void pwmhip_server::run_build(Worker* w, Job& job) {
  std::vector<const void*> texts;
  std::vector<uint64_t> sizes;
  uint32_t width = 1, levels = 0, algo = 0, flags = 0;
  {
    std::lock_guard<std::mutex> g(mu);
    for (uint64_t id : job.builds) {
      const Index& x = indices[id];
      texts.push_back(x.text);
      sizes.push_back(x.n);
      width = x.width;
      levels = x.levels;
      algo = x.algo;
      flags = x.flags;
    }
  }
  const uint64_t T = texts.size();
  std::vector<pwmhip_ws> outs(T);
  int st = pwmhip_build_batch(w->ctx, texts.data(), sizes.data(), T, width, levels, algo,
                              flags & ~uint32_t(PWMHIP_FLAG_REUSE_OUTPUT), outs.data());
  std::vector<pwmhip_rs*> rs(T, nullptr);
  std::vector<int> sts(T, st);
  if (st != PWMHIP_OK && T > 1) {
    // one bad request (say a symbol out of range) must not fail the others:
    // build them one by one, each structure owning its memory
    std::vector<pwmhip_ws> one(T);
    for (uint64_t t = 0; t < T; ++t) {
      sts[t] = pwmhip_build(w->ctx, texts[t], sizes[t], width, levels, algo,
                            flags & ~uint32_t(PWMHIP_FLAG_REUSE_OUTPUT), &one[t]);
      if (sts[t] == PWMHIP_OK && one[t].levels) sts[t] = pwmhip_rs_build(w->ctx, &one[t], &rs[t]);
    }
    outs = one;
    st = PWMHIP_ERR_INVALID;  // no shared block
  } else if (st == PWMHIP_OK)
    for (uint64_t t = 0; t < T; ++t)
      if (outs[t].levels) sts[t] = pwmhip_rs_build(w->ctx, &outs[t], &rs[t]);
  // scratch above the allowance goes back
  if (w->ctx->scratch.cap + w->ctx->text_scratch.cap + w->ctx->rs_scratch.cap > allowance) {
    (void)hipStreamSynchronize(w->ctx->stream);
    w->ctx->scratch.release();
    w->ctx->text_scratch.release();
  }
  const auto now = clk::now();
  std::lock_guard<std::mutex> g(mu);
  build_launches++;
  texts_launched += T;
  for (uint64_t id : job.builds) charge(-int64_t(indices[id].est));  // estimates out, actual bytes in
  std::shared_ptr<Block> block;
  if (st == PWMHIP_OK && T && outs[0].alloc_bytes && (T == 1 || outs[1].alloc_bytes == 0)) {
    block = std::make_shared<Block>();
    block->owner = outs[0];
    block->owner.code_len = block->owner.code_word = nullptr;
    block->owner.n_codes = 0;
    block->bytes = outs[0].alloc_bytes;
    outs[0].alloc_bytes = 0;
    charge(int64_t(block->bytes));
  }
  for (uint64_t t = 0; t < T; ++t) {
    Index& x = indices[job.builds[t]];
    if (sts[t] != PWMHIP_OK) {
      x.status = sts[t];
      builds_failed++;
      if (rs[t]) pwmhip_rs_free(w->ctx, rs[t]);
      if (outs[t].alloc_bytes) pwmhip_ws_free(w->ctx, &outs[t]);
      continue;
    }
    x.ws = outs[t];
    x.rs = rs[t];
    x.bytes = (x.ws.alloc_bytes) + (x.ws.levels ? rs_bytes(x.n, x.levels) : 0);
    charge(int64_t(x.bytes));
    if (block) {
      x.block = block;
      block->refs++;
    }
    x.status = 0;
    build_lat.push_back(ms_since(x.t_submit, now));
  }
  if (block && block->refs == 0) {  // every structure of the batch failed
    pwmhip_ws_free(w->ctx, &block->owner);
    charge(-int64_t(block->bytes));
  }
}

void pwmhip_server::run_queries(Worker* w, Job& job) {
  // distinct indices -> descriptor slots
  std::vector<pwm::QueryIndexDesc> desc;
  std::unordered_map<uint64_t, uint32_t> slot;
  uint64_t n = 0;
  {
    std::lock_guard<std::mutex> g(mu);
    for (auto& q : job.queries) {
      if (!slot.count(q.index)) {
        slot[q.index] = uint32_t(desc.size());
        const Index& x = indices[q.index];
        pwm::QueryIndexDesc d{};
        if (x.rs) d = pwm::query_desc(x.rs);
        desc.push_back(d);
      }
      n += q.n;
    }
  }
  int st = PWMHIP_OK;
  try {
    pwmhip_ctx* c = w->ctx;
    PWM_HIP_TRY(hipSetDevice(c->device));
    auto al = [](uint64_t x) { return (x + 255) & ~uint64_t(255); };
    const uint64_t need = al(desc.size() * sizeof(pwm::QueryIndexDesc)) + al(n * 4) * 2 + al(n * 8) * 3 + 4096;
    c->rs_scratch.ensure(need);
    c->rs_scratch.reset();
    auto* d_desc = c->rs_scratch.take<pwm::QueryIndexDesc>(desc.size());
    auto* d_idx = c->rs_scratch.take<uint32_t>(n);
    auto* d_op = c->rs_scratch.take<uint32_t>(n);
    auto* d_a = c->rs_scratch.take<uint64_t>(n);
    auto* d_b = c->rs_scratch.take<uint64_t>(n);
    auto* d_r = c->rs_scratch.take<uint64_t>(n);
    // pack on the host, one copy per array
    std::vector<uint32_t> h_idx(n), h_op(n);
    std::vector<uint64_t> h_a(n), h_b(n), h_r(n);
    uint64_t o = 0;
    for (auto& q : job.queries) {
      const uint32_t sl = slot[q.index];
      std::fill(h_idx.begin() + o, h_idx.begin() + o + q.n, sl);
      std::memcpy(h_op.data() + o, q.op, q.n * 4);
      std::memcpy(h_a.data() + o, q.a, q.n * 8);
      if (q.b) std::memcpy(h_b.data() + o, q.b, q.n * 8);
      else std::memcpy(h_b.data() + o, q.a, q.n * 8);
      o += q.n;
    }
    PWM_HIP_TRY(hipMemcpyAsync(d_desc, desc.data(), desc.size() * sizeof(pwm::QueryIndexDesc),
                               hipMemcpyHostToDevice, c->stream));
    PWM_HIP_TRY(hipMemcpyAsync(d_idx, h_idx.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    PWM_HIP_TRY(hipMemcpyAsync(d_op, h_op.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    PWM_HIP_TRY(hipMemcpyAsync(d_a, h_a.data(), n * 8, hipMemcpyHostToDevice, c->stream));
    PWM_HIP_TRY(hipMemcpyAsync(d_b, h_b.data(), n * 8, hipMemcpyHostToDevice, c->stream));
    pwm::launch_multi_query(c->stream, d_desc, d_idx, d_op, d_a, d_b, d_r, n);
    PWM_HIP_TRY(hipMemcpyAsync(h_r.data(), d_r, n * 8, hipMemcpyDeviceToHost, c->stream));
    PWM_HIP_TRY(hipStreamSynchronize(c->stream));
    o = 0;
    for (auto& q : job.queries) {
      std::memcpy(q.result, h_r.data() + o, q.n * 8);
      o += q.n;
    }
  } catch (const pwm::hip_failure& f) {
    (void)hipGetLastError();
    st = f.status;
  } catch (const pwm::status_error& e) {
    st = e.status;
  } catch (...) {
    st = PWMHIP_ERR_OOM;
  }
  const auto now = clk::now();
  std::lock_guard<std::mutex> g(mu);
  query_launches++;
  for (auto& q : job.queries) {
    query_status[q.id] = st;
    queries_done += q.n;
    query_lat.push_back(ms_since(q.t_submit, now));
    auto ix = indices.find(q.index);
    if (ix != indices.end()) {
      ix->second.queries_pending--;
      if (ix->second.release_requested && ix->second.queries_pending == 0) drop_index(q.index);
    }
  }
}

// A job whose host bookkeeping threw (an allocation): requests it left
// pending fail with `st`.  Under mu.
void pwmhip_server::fail_job(const Job& job, int st) {
  for (uint64_t id : job.builds) {
    auto it = indices.find(id);
    if (it == indices.end() || it->second.status != 1) continue;
    it->second.status = st;
    builds_failed++;
    charge(-int64_t(it->second.est));
  }
  for (const auto& q : job.queries) {
    auto qs = query_status.find(q.id);
    if (qs == query_status.end() || qs->second != 1) continue;
    qs->second = st;
    auto ix = indices.find(q.index);
    if (ix != indices.end() && ix->second.queries_pending) ix->second.queries_pending--;
  }
}

void pwmhip_server::worker(Worker* w) {
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    cv_work.wait(lk, [&] { return w->busy || exiting; });
    if (!w->busy) return;
    Job job = std::move(w->job);
    lk.unlock();
    const int st = pwm::guarded([&] {
      if (job.kind == Job::BUILD) run_build(w, job);
      else run_queries(w, job);
    });
    lk.lock();
    if (st != PWMHIP_OK) fail_job(job, st);
    w->busy = false;
    cv_sched.notify_all();
    cv_done.notify_all();
  }
}

extern "C" {

int pwmhip_server_create(const pwmhip_server_config* cfg, pwmhip_server** out) {
  if (!out) return PWMHIP_ERR_INVALID;
  *out = nullptr;
  auto* s = new (std::nothrow) pwmhip_server();
  if (!s) return PWMHIP_ERR_OOM;
  if (cfg) s->cfg = *cfg;
  if (!s->cfg.hbm_budget_bytes) s->cfg.hbm_budget_bytes = 200ull * 1000 * 1000 * 1000;
  if (!s->cfg.workers) s->cfg.workers = 2;
  if (!s->cfg.max_batch_texts) s->cfg.max_batch_texts = 1024;
  if (!s->cfg.max_batch_bytes) s->cfg.max_batch_bytes = 512ull << 20;
  if (!s->cfg.max_query_batch) s->cfg.max_query_batch = 8ull << 20;
  const uint64_t mb = s->cfg.max_batch_bytes;
  // scratch of the largest batch (8 levels) plus query staging
  s->allowance = scratch_bytes((mb + kChunk - 1) / kChunk + s->cfg.max_batch_texts, s->cfg.max_batch_texts, mb, 8) +
                 s->cfg.max_query_batch * 48 + (64ull << 20);
  if (uint64_t(s->cfg.workers) * s->allowance >= s->cfg.hbm_budget_bytes) {
    delete s;
    return PWMHIP_ERR_BUDGET;
  }
  int st = pwm::guarded([&] {
    for (uint32_t i = 0; i < s->cfg.workers; ++i) {
      auto w = std::make_unique<pwmhip_server::Worker>();
      const int cst = pwmhip_create(&w->ctx, s->cfg.device);
      if (cst != PWMHIP_OK) throw pwm::status_error{cst, "create"};
      s->workers.push_back(std::move(w));
    }
  });
  if (st != PWMHIP_OK) {
    for (auto& x : s->workers) pwmhip_destroy(x->ctx);
    delete s;
    return st;
  }
  s->charge(int64_t(s->cfg.workers) * int64_t(s->allowance));
  // threads: a failed start stops and joins the ones already running
  st = pwm::guarded([&] {
    for (auto& w : s->workers) w->th = std::thread(&pwmhip_server::worker, s, w.get());
    s->sched = std::thread(&pwmhip_server::scheduler, s);
  });
  if (st != PWMHIP_OK) {
    {
      std::lock_guard<std::mutex> g(s->mu);
      s->exiting = true;
    }
    s->cv_work.notify_all();
    for (auto& w : s->workers)
      if (w->th.joinable()) w->th.join();
    for (auto& w : s->workers) pwmhip_destroy(w->ctx);
    delete s;
    return st;
  }
  *out = s;
  return PWMHIP_OK;
}

int pwmhip_server_destroy(pwmhip_server* s) {
  if (!s) return PWMHIP_ERR_INVALID;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_sched.notify_all();
  s->cv_work.notify_all();
  s->sched.join();
  for (auto& w : s->workers) w->th.join();
  {
    std::lock_guard<std::mutex> g(s->mu);
    while (!s->indices.empty()) s->drop_index(s->indices.begin()->first);
  }
  for (auto& w : s->workers) pwmhip_destroy(w->ctx);
  delete s;
  return PWMHIP_OK;
}

int pwmhip_server_build(pwmhip_server* s, const void* text, uint64_t size, uint32_t word_width, uint32_t levels,
                        uint32_t algo, uint32_t flags, uint64_t* index_id) {
  if (!s || !index_id || (size && !text)) return PWMHIP_ERR_INVALID;
  if (word_width != 1 && word_width != 2 && word_width != 4 && word_width != 8) return PWMHIP_ERR_INVALID;
  if (algo > PWMHIP_ALGO_NAIVE) return PWMHIP_ERR_UNSUPPORTED;  // plain shapes (sizes known at admission)
  if (levels == 0 || levels > 8 * word_width || levels > PWMHIP_MAX_LEVELS) return PWMHIP_ERR_INVALID;
  if (flags & (PWMHIP_FLAG_TEXT_ON_DEVICE | PWMHIP_FLAG_REUSE_OUTPUT)) return PWMHIP_ERR_INVALID;
  Index x;
  x.n = size;
  x.width = word_width;
  x.levels = levels;
  x.algo = algo;
  x.flags = flags;
  x.text = text;
  x.est = output_bytes(size, levels) + rs_bytes(size, levels) + 4096;
  x.t_submit = clk::now();
  const int st = pwm::guarded([&] {
    std::lock_guard<std::mutex> g(s->mu);
    const uint64_t fixed = uint64_t(s->cfg.workers) * s->allowance;
    const uint64_t chunks = (size + kChunk - 1) / kChunk;
    if (fixed + x.est > s->cfg.hbm_budget_bytes ||
        (levels <= pwm::kMaxGroup && scratch_bytes(chunks, 1, size * word_width, levels) > s->allowance &&
         size * word_width > s->cfg.max_batch_bytes))
      throw pwm::status_error{PWMHIP_ERR_BUDGET, "build"};
    const uint64_t id = s->next_id;
    s->indices.emplace(id, x);
    try {
      s->pending_builds.push_back(id);
    } catch (...) {
      s->indices.erase(id);
      throw;
    }
    s->next_id++;
    *index_id = id;
  });
  if (st != PWMHIP_OK) return st;
  s->cv_sched.notify_all();
  return PWMHIP_OK;
}

int pwmhip_server_query(pwmhip_server* s, uint64_t index_id, const uint32_t* op, const uint64_t* a,
                        const uint64_t* b, uint64_t n, uint64_t* result, uint64_t* request_id) {
  if (!s || !request_id || (n && (!op || !a || !result))) return PWMHIP_ERR_INVALID;
  QueryReq q;
  q.index = index_id;
  q.op = op;
  q.a = a;
  q.b = b;
  q.n = n;
  q.result = result;
  q.t_submit = clk::now();
  bool queued = false;
  const int st = pwm::guarded([&] {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->indices.find(index_id);
    if (it == s->indices.end() || it->second.release_requested) throw pwm::status_error{PWMHIP_ERR_INVALID, "query"};
    if (n > s->cfg.max_query_batch) throw pwm::status_error{PWMHIP_ERR_BUDGET, "query"};
    q.id = s->next_id;
    if (n == 0) {
      s->query_status[q.id] = PWMHIP_OK;
    } else {
      s->query_status[q.id] = 1;
      try {
        s->pending_queries.push_back(q);
      } catch (...) {
        s->query_status.erase(q.id);
        throw;
      }
      it->second.queries_pending++;
      queued = true;
    }
    s->next_id++;
    *request_id = q.id;
  });
  if (st != PWMHIP_OK) return st;
  if (queued) s->cv_sched.notify_all();
  return PWMHIP_OK;
}

int pwmhip_server_wait(pwmhip_server* s, uint64_t id, int64_t timeout_ms) {
  if (!s) return PWMHIP_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  int result = PWMHIP_ERR_INVALID;
  auto done = [&] {
    auto q = s->query_status.find(id);
    if (q != s->query_status.end()) {
      if (q->second == 1) return false;
      result = q->second;
      return true;
    }
    auto ix = s->indices.find(id);
    if (ix != s->indices.end()) {
      if (ix->second.status == 1) return false;
      result = ix->second.status;
      return true;
    }
    result = PWMHIP_ERR_INVALID;
    return true;
  };
  if (timeout_ms < 0) {
    s->cv_done.wait(lk, done);
    return result;
  }
  if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done)) return PWMHIP_ERR_TIMEOUT;
  return result;
}

int pwmhip_server_index_info(pwmhip_server* s, uint64_t index_id, pwmhip_ws* info) {
  if (!s || !info) return PWMHIP_ERR_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->indices.find(index_id);
  if (it == s->indices.end() || it->second.status != 0) return PWMHIP_ERR_INVALID;
  *info = it->second.ws;
  info->alloc_bytes = 0;  // owned by the server
  info->code_len = info->code_word = nullptr;
  info->n_codes = 0;
  return PWMHIP_OK;
}

int pwmhip_server_download(pwmhip_server* s, uint64_t index_id, uint64_t level, uint64_t* host_words) {
  if (!s || !host_words) return PWMHIP_ERR_INVALID;
  pwmhip_ws ws{};
  pwmhip_ctx* c = nullptr;
  {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->indices.find(index_id);
    if (it == s->indices.end() || it->second.status != 0 || level >= it->second.ws.levels) return PWMHIP_ERR_INVALID;
    ws = it->second.ws;
    c = s->workers[0]->ctx;
  }
  // a synchronous copy (the index is complete: its worker synchronised)
  const uint64_t nw = pwm::words_for(ws.level_bits[level]);
  if (!nw) return PWMHIP_OK;
  if (hipSetDevice(c->device) != hipSuccess) return PWMHIP_ERR_HIP;
  if (hipMemcpy(host_words, ws.d_words + ws.level_offset[level], nw * 8, hipMemcpyDeviceToHost) != hipSuccess)
    return PWMHIP_ERR_HIP;
  return PWMHIP_OK;
}

int pwmhip_server_release(pwmhip_server* s, uint64_t index_id) {
  if (!s) return PWMHIP_ERR_INVALID;
  {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->indices.find(index_id);
    if (it == s->indices.end()) return PWMHIP_ERR_INVALID;
    if (it->second.status == 1) return PWMHIP_ERR_INVALID;  // still building: wait first
    if (it->second.queries_pending) it->second.release_requested = true;
    else s->drop_index(index_id);
  }
  s->cv_sched.notify_all();
  return PWMHIP_OK;
}

int pwmhip_server_stats_get(pwmhip_server* s, pwmhip_server_stats* out) {
  if (!s || !out) return PWMHIP_ERR_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  std::memset(out, 0, sizeof(*out));
  return pwm::guarded([&] {
  auto pct = [](std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const size_t k = std::min(v.size() - 1, size_t(p * double(v.size() - 1) + 0.5));
    return v[k];
  };
  out->builds_done = s->build_lat.size();
  out->builds_failed = s->builds_failed;
  out->queries_done = s->queries_done;
  out->query_requests_done = s->query_lat.size();
  out->build_launches = s->build_launches;
  out->query_launches = s->query_launches;
  out->build_p50_ms = pct(s->build_lat, 0.5);
  out->build_p99_ms = pct(s->build_lat, 0.99);
  out->query_p50_ms = pct(s->query_lat, 0.5);
  out->query_p99_ms = pct(s->query_lat, 0.99);
  out->bytes_in_use = s->in_use;
  out->bytes_peak = s->peak;
  out->mean_texts_per_launch = s->build_launches ? double(s->texts_launched) / double(s->build_launches) : 0.0;
  });
}

}  // extern "C"
