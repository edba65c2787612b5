// y3g_server.cpp -- request-serving front end of liby3g (y3g_server_* in
// include/y3g.h).
//
// The reference evaluates one CosmoSIS sample at a time, each execute()
// looping over its integrals (src/utils/CosmoSISSICUDAModule.hh:282-335).
// A sampler that proposes many parameter points -- or several chains
// sharing one GPU -- instead submits requests here:
//
//   request   = one parameter point + the integrals wanted at it;
//   lanes     = independent runtimes (own HIP stream, tables, region slab),
//               each running one batch at a time, so a new batch starts as
//               soon as any lane is free (continuous batching at batch
//               granularity) and small batches share the device;
//   admission = the lane count is capped so that every lane's region slab
//               fits the HBM budget; requests queue (FIFO) until a lane
//               takes them, whole, up to max_batch_integrals per batch;
//   cache     = an LRU map from (parameter point, integral) to its result,
//               so a repeated request is answered without the device.
//
// Per-batch tables are rebuilt on the lane's device (a fraction of a
// millisecond per parameter point), so the cache keeps results rather
// than tables.  Host code only: the device work is y3g_tables_build and
// y3g_integrate on the lane's runtime.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "y3g.h"

namespace {

thread_local std::string g_srv_err;

int srv_fail(int code, std::string const& m)
{
  g_srv_err = m;
  return code;
}

double now_s()
{
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
  unsigned long long ticket = 0;
  y3g_params params{};
  std::vector<y3g_integral> igs;   // the integrals still to compute
  std::vector<size_t> slot;        // their positions in the caller's array
  y3g_result* out = nullptr;
  double t_submit = 0.0, t_done = 0.0;
  int state = 0;                   // 0 queued, 1 running, 2 done
  int rc = 0;
  std::string msg;
};

std::string cache_key(const y3g_params& p, const y3g_integral& g)
{
  y3g_integral k = g;
  k.point = 0;
  std::string s(reinterpret_cast<const char*>(&p), sizeof(p));
  s.append(reinterpret_cast<const char*>(&k), sizeof(k));
  return s;
}

}  // namespace

struct y3g_server {
  y3g_server_options opt{};
  std::vector<y3g_runtime*> rts;
  std::vector<std::thread> threads;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::unordered_map<unsigned long long, std::shared_ptr<Request>> live;
  unsigned long long next_ticket = 1;
  bool stop = false;
  // LRU result cache
  std::list<std::string> lru;
  std::unordered_map<std::string, std::pair<y3g_result, std::list<std::string>::iterator>> cache;
  y3g_server_stats stats{};

  void cache_put(const std::string& k, const y3g_result& r)
  {
    if (opt.cache_entries <= 0) return;
    auto it = cache.find(k);
    if (it != cache.end()) {
      lru.erase(it->second.second);
      cache.erase(it);
    }
    lru.push_front(k);
    cache.emplace(k, std::make_pair(r, lru.begin()));
    while ((long long)cache.size() > opt.cache_entries) {
      cache.erase(lru.back());
      lru.pop_back();
    }
  }

  bool cache_get(const std::string& k, y3g_result* r)
  {
    auto it = cache.find(k);
    if (it == cache.end()) return false;
    lru.splice(lru.begin(), lru, it->second.second);
    *r = it->second.first;
    return true;
  }

  void lane_loop(int lane)
  {
    y3g_runtime* rt = rts[lane];
    std::vector<std::shared_ptr<Request>> batch;
    std::vector<y3g_integral> igs;
    std::vector<y3g_result> res;
    std::vector<y3g_params> pts;
    while (true) {
      batch.clear();
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || !queue.empty(); });
        if (queue.empty()) return;  // stop requested and nothing left
        size_t n = 0;
        while (!queue.empty()) {
          auto const& r = queue.front();
          if (!batch.empty() && n + r->igs.size() > (size_t)opt.max_batch_integrals) break;
          n += r->igs.size();
          r->state = 1;
          batch.push_back(r);
          queue.pop_front();
        }
      }
      // one table set per distinct parameter point of the batch
      pts.clear();
      igs.clear();
      std::map<std::string, int> pidx;
      for (auto& r : batch) {
        std::string k(reinterpret_cast<const char*>(&r->params), sizeof(y3g_params));
        auto it = pidx.find(k);
        int p;
        if (it == pidx.end()) {
          p = (int)pts.size();
          pidx.emplace(k, p);
          pts.push_back(r->params);
        } else {
          p = it->second;
        }
        for (auto g : r->igs) {
          g.point = p;
          igs.push_back(g);
        }
      }
      res.assign(igs.size(), y3g_result{});
      int rc = y3g_tables_build(rt, pts.data(), &opt.spec, (int)pts.size());
      if (rc == Y3G_OK) rc = y3g_integrate(rt, igs.data(), igs.size(), &opt.options, res.data());
      std::string const msg = rc == Y3G_OK ? std::string() : std::string(y3g_last_error());
      double kernel_ms = 0.0;
      long long nev = 0;
      if (rc == Y3G_OK) y3g_last_timing(rt, &kernel_ms, &nev);
      {
        std::lock_guard<std::mutex> lk(mu);
        size_t off = 0;
        double const t = now_s();
        for (auto& r : batch) {
          for (size_t i = 0; i < r->igs.size(); ++i) {
            if (rc == Y3G_OK) {
              r->out[r->slot[i]] = res[off + i];
              cache_put(cache_key(r->params, r->igs[i]), res[off + i]);
            }
          }
          off += r->igs.size();
          r->rc = rc;
          r->msg = msg;
          r->state = 2;
          r->t_done = t;
          stats.requests_done += 1;
          stats.integrals_computed += (long long)r->igs.size();
        }
        stats.batches += 1;
        stats.kernel_ms += kernel_ms;
        stats.evaluations += nev;
      }
      cv_done.notify_all();
    }
  }
};

extern "C" {

const char* y3g_server_last_error(void) { return g_srv_err.c_str(); }

void y3g_server_options_default(y3g_server_options* o)
{
  std::memset(o, 0, sizeof(*o));
  o->lanes = 2;
  o->max_batch_integrals = 4096;
  o->hbm_budget_bytes = (size_t)64 << 30;
  o->cache_entries = 1 << 20;
  y3g_options_default(&o->options);
  y3g_spec_wide(&o->spec);
}

int y3g_server_create(int device, const y3g_server_options* opt, y3g_server** out)
{
  if (!out) return srv_fail(Y3G_ERR_ARG, "y3g_server_create: null out");
  *out = nullptr;
  auto s = std::make_unique<y3g_server>();
  if (opt)
    s->opt = *opt;
  else
    y3g_server_options_default(&s->opt);
  if (s->opt.lanes < 1 || s->opt.max_batch_integrals < 1)
    return srv_fail(Y3G_ERR_ARG, "y3g_server_create: lanes and max_batch_integrals must be >= 1");
  // admission: as many lanes as the region slabs fit in the budget
  for (int l = 0; l < s->opt.lanes; ++l) {
    y3g_runtime* rt = nullptr;
    if (y3g_runtime_create(device, &rt) != Y3G_OK) {
      for (auto* r : s->rts) y3g_runtime_destroy(r);
      return srv_fail(Y3G_ERR_NODEVICE, std::string("y3g_server_create: ") + y3g_last_error());
    }
    size_t bytes = 0;
    y3g_runtime_footprint(rt, &s->opt.options, &bytes);
    if ((l + 1) * bytes > s->opt.hbm_budget_bytes) {
      y3g_runtime_destroy(rt);
      if (l == 0)
        return srv_fail(Y3G_ERR_MEMORY, "y3g_server_create: one lane needs " +
                                            std::to_string(bytes) + " bytes, over the budget");
      break;
    }
    s->rts.push_back(rt);
    s->stats.slab_bytes_per_lane = (long long)bytes;
  }
  s->stats.lanes = (int)s->rts.size();
  y3g_server* raw = s.release();
  for (int l = 0; l < (int)raw->rts.size(); ++l)
    raw->threads.emplace_back([raw, l] { raw->lane_loop(l); });
  *out = raw;
  return Y3G_OK;
}

int y3g_server_submit(y3g_server* s, const y3g_params* params, const y3g_integral* igs,
                      size_t n, y3g_result* out, unsigned long long* ticket)
{
  if (!s || !params || !ticket || (n && (!igs || !out)))
    return srv_fail(Y3G_ERR_ARG, "y3g_server_submit: bad arguments");
  for (size_t i = 0; i < n; ++i)
    if (igs[i].ndim < 2 || igs[i].ndim > 6 || igs[i].kind < 0 || igs[i].kind > 2)
      return srv_fail(Y3G_ERR_ARG, "y3g_server_submit: integral " + std::to_string(i) +
                                       " is malformed");
  auto r = std::make_shared<Request>();
  r->params = *params;
  r->out = out;
  r->t_submit = now_s();
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return srv_fail(Y3G_ERR_STATE, "y3g_server_submit: server is stopping");
  for (size_t i = 0; i < n; ++i) {
    y3g_result hit;
    if (s->opt.cache_entries > 0 && s->cache_get(cache_key(*params, igs[i]), &hit)) {
      out[i] = hit;
      s->stats.cache_hits += 1;
    } else {
      r->igs.push_back(igs[i]);
      r->slot.push_back(i);
    }
  }
  r->ticket = s->next_ticket++;
  *ticket = r->ticket;
  s->live.emplace(r->ticket, r);
  s->stats.requests_submitted += 1;
  if (r->igs.empty()) {
    r->state = 2;
    r->t_done = now_s();
    s->stats.requests_done += 1;
    s->cv_done.notify_all();
  } else {
    s->queue.push_back(r);
    s->cv_work.notify_one();
  }
  return Y3G_OK;
}

int y3g_server_wait(y3g_server* s, unsigned long long ticket, double timeout_s,
                    double* latency_s)
{
  if (!s) return srv_fail(Y3G_ERR_ARG, "y3g_server_wait: null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->live.find(ticket);
  if (it == s->live.end())
    return srv_fail(Y3G_ERR_ARG, "y3g_server_wait: unknown ticket " + std::to_string(ticket));
  std::shared_ptr<Request> r = it->second;
  auto done = [&] { return r->state == 2; };
  if (timeout_s < 0)
    s->cv_done.wait(lk, done);
  else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done))
    return srv_fail(Y3G_ERR_STATE, "y3g_server_wait: timed out");
  s->live.erase(ticket);
  if (latency_s) *latency_s = r->t_done - r->t_submit;
  if (r->rc != Y3G_OK) return srv_fail(r->rc, "y3g_server: batch failed: " + r->msg);
  return Y3G_OK;
}

int y3g_server_stats_get(y3g_server* s, y3g_server_stats* st)
{
  if (!s || !st) return srv_fail(Y3G_ERR_ARG, "y3g_server_stats_get: bad arguments");
  std::lock_guard<std::mutex> lk(s->mu);
  *st = s->stats;
  st->queued = (long long)s->queue.size();
  return Y3G_OK;
}

int y3g_server_destroy(y3g_server* s)
{
  if (!s) return Y3G_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& t : s->threads) t.join();  // lanes finish the queued work first
  for (auto* rt : s->rts) y3g_runtime_destroy(rt);
  delete s;
  return Y3G_OK;
}

}  // extern "C"
