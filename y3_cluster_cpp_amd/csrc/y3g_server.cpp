// y3g_server.cpp -- request-serving front end of liby3g (y3g_server_* in
// include/y3g.h).
//
// The reference evaluates one CosmoSIS sample at a time, each execute()
// looping over its integrals (src/utils/CosmoSISSICUDAModule.hh:282-335).
// A sampler that proposes many parameter points -- or several chains
// sharing one GPU -- instead submits requests here:
//
//   request   = one parameter point + the integrals wanted at it;
//   launch    = ONE persistent streaming cubature launch (y3g_stream.hpp)
//               that takes integrals from a ring in pinned host memory as
//               they are published, so a request that arrives while others
//               are on the GPU joins the running launch and its integrals
//               start as soon as workgroups free up (continuous batching);
//   tables    = per-cosmology table slots on the device under an LRU: a
//               repeated parameter point reuses its slot, a new one takes a
//               free or least-recently-used idle slot and its table build is
//               published ahead of its integrals, into the same launch;
//   admission = the session's region slabs and table_slots table sets must
//               fit hbm_budget_bytes (fewer slots are kept otherwise); a
//               request is admitted, in arrival order, once it has a table
//               slot and ring room for its integrals;
//   cache     = an LRU map from (parameter point, integral) to its result,
//               so a repeated request is answered without the device.
//
// A poller thread collects finished integrals from the ring, completes
// requests, admits queued ones and starts a new launch when published work
// is left and the previous launch has ended (an idle launch ends by itself).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "y3g.h"
#include "y3g_stream.hpp"

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
  std::string key;                 // the parameter point's bytes
  std::vector<y3g_integral> igs;   // the integrals still to compute
  std::vector<size_t> slot;        // their positions in the caller's array
  y3g_result* out = nullptr;
  size_t remaining = 0;
  int table_slot = -1;
  double t_submit = 0.0, t_done = 0.0;
  int state = 0;                   // 0 queued, 1 in flight, 2 done
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

int gm_nodes(int d) { return 1 + 4 * d + 2 * d * (d - 1) + (1 << d); }

}  // namespace

struct y3g_server {
  y3g_server_options opt{};
  y3g_runtime* rt = nullptr;
  y3g_stream* st = nullptr;
  std::thread poller;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::unordered_map<unsigned long long, std::shared_ptr<Request>> live;
  unsigned long long next_ticket = 1;
  bool stop = false;
  // ring entry (modulo the ring) -> its request and position
  std::vector<std::pair<std::shared_ptr<Request>, size_t>> inflight;
  long long n_inflight = 0;
  // table slots: key, generation, integrals in flight, LRU order
  std::vector<std::string> slot_key;
  std::vector<unsigned> slot_gen;
  std::vector<long long> slot_pending;
  std::list<int> slot_lru;  // front = most recently used
  std::vector<std::list<int>::iterator> slot_pos;
  std::unordered_map<std::string, int> slot_of;
  // LRU result cache
  std::list<std::string> lru;
  std::unordered_map<std::string, std::pair<y3g_result, std::list<std::string>::iterator>> cache;
  y3g_server_stats stats{};
  std::string fatal;  // a device error ends the session

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

  void finish(const std::shared_ptr<Request>& r, int rc, std::string const& msg)
  {
    r->rc = rc;
    r->msg = msg;
    r->state = 2;
    r->t_done = now_s();
    stats.requests_done += 1;
    cv_done.notify_all();
  }

  // A slot for key: its own if resident, else a never-used or the least
  // recently used slot with nothing in flight (-1: none now).
  int slot_for(const std::string& key, bool* build)
  {
    auto it = slot_of.find(key);
    if (it != slot_of.end()) {
      *build = false;
      return it->second;
    }
    *build = true;
    for (auto p = slot_lru.rbegin(); p != slot_lru.rend(); ++p) {
      int const k = *p;
      if (slot_pending[k] == 0 && y3g::stream_tables_pending(st, k) == 0) return k;
    }
    return -1;
  }

  // Publish queued requests, in arrival order, while slots and ring room last.
  void admit()
  {
    while (!queue.empty() && fatal.empty()) {
      auto r = queue.front();
      bool build = false;
      int const k = slot_for(r->key, &build);
      if (k < 0) return;
      size_t const need = r->igs.size() + (build ? 1 : 0);
      if (y3g::stream_room(st) < need) return;
      if (build) {
        if (!slot_key[k].empty()) slot_of.erase(slot_key[k]);
        unsigned g = 0;
        if (y3g::stream_tables(st, k, r->params, &g) != Y3G_OK) {
          fatal = y3g_last_error();
          return;
        }
        slot_key[k] = r->key;
        slot_gen[k] = g;
        slot_of[r->key] = k;
        stats.table_builds += 1;
      } else {
        stats.table_reuses += 1;
      }
      unsigned long long first = 0;
      if (y3g::stream_push(st, k, slot_gen[k], r->igs.data(), r->igs.size(), &first) !=
          Y3G_OK) {
        queue.pop_front();
        finish(r, Y3G_ERR_ARG, y3g_last_error());
        continue;
      }
      size_t const cap = inflight.size();
      for (size_t i = 0; i < r->igs.size(); ++i)
        inflight[(size_t)((first + i) % cap)] = {r, i};
      n_inflight += (long long)r->igs.size();
      slot_pending[k] += (long long)r->igs.size();
      slot_lru.splice(slot_lru.begin(), slot_lru, slot_pos[k]);
      r->table_slot = k;
      r->state = 1;
      queue.pop_front();
    }
  }

  static void on_result(void* ctx, unsigned long long item, const y3g_result& res)
  {
    auto* s = static_cast<y3g_server*>(ctx);
    auto& e = s->inflight[(size_t)(item % s->inflight.size())];
    std::shared_ptr<Request> r = std::move(e.first);
    size_t const i = e.second;
    r->out[r->slot[i]] = res;
    if (res.status != Y3G_STATUS_NO_TABLES) s->cache_put(cache_key(r->params, r->igs[i]), res);
    s->slot_pending[r->table_slot] -= 1;
    s->n_inflight -= 1;
    s->stats.integrals_computed += 1;
    s->stats.evaluations += res.neval;
    if (--r->remaining == 0) s->finish(r, Y3G_OK, std::string());
  }

  void fail_all(std::string const& msg)
  {
    for (auto& e : inflight)
      if (e.first && e.first->state != 2) finish(e.first, Y3G_ERR_HIP, msg);
    for (auto& r : queue) finish(r, Y3G_ERR_HIP, msg);
    queue.clear();
    for (auto& e : inflight) e.first.reset();
    n_inflight = 0;
  }

  void poll_loop()
  {
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
      if (fatal.empty()) {
        size_t const got = y3g::stream_collect(st, &y3g_server::on_result, this);
        admit();
        if (fatal.empty() && y3g::stream_kick(st) != Y3G_OK) fatal = y3g_last_error();
        if (!fatal.empty()) fail_all("y3g_server: device error: " + fatal);
        if (got) continue;
      }
      if (stop && queue.empty() && n_inflight == 0) return;
      if (n_inflight == 0 && queue.empty())
        cv_work.wait(lk);  // idle: a submit or the stop wakes it
      else
        cv_work.wait_for(lk, std::chrono::microseconds(20));
    }
  }
};

extern "C" {

const char* y3g_server_last_error(void) { return g_srv_err.c_str(); }

void y3g_server_options_default(y3g_server_options* o)
{
  std::memset(o, 0, sizeof(*o));
  o->ring_integrals = 65536;
  o->table_slots = 256;
  o->hbm_budget_bytes = (size_t)64 << 30;
  o->cache_entries = 1 << 20;
  o->idle_ms = 50.0;
  o->table_wait_ms = 2000.0;
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
  if (s->opt.ring_integrals < 2 || s->opt.table_slots < 1)
    return srv_fail(Y3G_ERR_ARG, "y3g_server_create: ring_integrals >= 2 and "
                                 "table_slots >= 1 are needed");
  if (y3g_runtime_create(device, &s->rt) != Y3G_OK)
    return srv_fail(Y3G_ERR_NODEVICE, std::string("y3g_server_create: ") + y3g_last_error());
  y3g::StreamConfig c{};
  c.ring_items = s->opt.ring_integrals;
  c.table_slots = s->opt.table_slots;
  c.idle_ms = s->opt.idle_ms;
  c.table_wait_ms = s->opt.table_wait_ms;
  c.options = s->opt.options;
  c.spec = s->opt.spec;
  // admission: the resident workgroups' region slabs, then as many table
  // slots as the rest of the budget holds (at most table_slots)
  size_t slab = 0, per_slot = 0;
  int wg = 0;
  auto bail = [&](int rc, std::string const& m) {
    y3g_runtime_destroy(s->rt);
    return srv_fail(rc, m);
  };
  if (y3g::stream_footprint(s->rt, c, &slab, &per_slot, &wg) != Y3G_OK)
    return bail(Y3G_ERR_ARG, std::string("y3g_server_create: ") + y3g_last_error());
  if (slab + per_slot > s->opt.hbm_budget_bytes)
    return bail(Y3G_ERR_MEMORY, "y3g_server_create: the region slabs and one table set need " +
                                  std::to_string(slab + per_slot) + " bytes, over the budget");
  size_t const fit = (s->opt.hbm_budget_bytes - slab) / per_slot;
  c.table_slots = (int)std::min<size_t>((size_t)c.table_slots, fit);
  if (y3g::stream_open(s->rt, c, &s->st) != Y3G_OK)
    return bail(Y3G_ERR_MEMORY, std::string("y3g_server_create: ") + y3g_last_error());
  s->inflight.resize((size_t)c.ring_items);
  s->slot_key.assign(c.table_slots, std::string());
  s->slot_gen.assign(c.table_slots, 0u);
  s->slot_pending.assign(c.table_slots, 0);
  for (int k = 0; k < c.table_slots; ++k) s->slot_lru.push_back(k);
  for (auto it = s->slot_lru.begin(); it != s->slot_lru.end(); ++it) s->slot_pos.push_back(it);
  s->stats.slab_bytes = (long long)slab;
  s->stats.table_bytes = (long long)(per_slot * (size_t)c.table_slots);
  s->stats.table_slots = c.table_slots;
  s->stats.workgroups = wg;
  y3g_server* raw = s.release();
  raw->poller = std::thread([raw] { raw->poll_loop(); });
  *out = raw;
  return Y3G_OK;
}

int y3g_server_submit(y3g_server* s, const y3g_params* params, const y3g_integral* igs,
                      size_t n, y3g_result* out, unsigned long long* ticket)
{
  if (!s || !params || !ticket || (n && (!igs || !out)))
    return srv_fail(Y3G_ERR_ARG, "y3g_server_submit: bad arguments");
  if (n + 1 > (size_t)s->opt.ring_integrals)
    return srv_fail(Y3G_ERR_ARG, "y3g_server_submit: a request of " + std::to_string(n) +
                                     " integrals is larger than the ring");
  for (size_t i = 0; i < n; ++i)
    if (igs[i].ndim < 2 || igs[i].ndim > 6 || igs[i].kind < 0 || igs[i].kind > 2 ||
        igs[i].algorithm != Y3G_PAGANI)
      return srv_fail(Y3G_ERR_ARG, "y3g_server_submit: integral " + std::to_string(i) +
                                       " is malformed (cubature, kinds 0-2)");
  auto r = std::make_shared<Request>();
  r->params = *params;
  r->key.assign(reinterpret_cast<const char*>(params), sizeof(*params));
  r->out = out;
  r->t_submit = now_s();
  y3g_options const& o = s->opt.options;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return srv_fail(Y3G_ERR_STATE, "y3g_server_submit: server is stopping");
  if (!s->fatal.empty())
    return srv_fail(Y3G_ERR_HIP, "y3g_server_submit: device error: " + s->fatal);
  for (size_t i = 0; i < n; ++i) {
    y3g_result hit;
    if (s->opt.cache_entries > 0 && s->cache_get(cache_key(*params, igs[i]), &hit)) {
      out[i] = hit;
      s->stats.cache_hits += 1;
      continue;
    }
    y3g_integral g = igs[i];
    // the tolerance bands' region budgets, as y3g_integrate applies them
    for (int k = 0; k < o.n_bands; ++k)
      if (g.eps_rel <= o.band_eps[k]) {
        long long const lim = 2LL * o.band_regions[k] * gm_nodes(g.ndim);
        if (lim < g.max_eval) g.max_eval = lim;
        break;
      }
    r->igs.push_back(g);
    r->slot.push_back(i);
  }
  r->remaining = r->igs.size();
  r->ticket = s->next_ticket++;
  *ticket = r->ticket;
  s->live.emplace(r->ticket, r);
  s->stats.requests_submitted += 1;
  if (r->igs.empty()) {
    s->finish(r, Y3G_OK, std::string());
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
  if (r->rc != Y3G_OK) return srv_fail(r->rc, "y3g_server: request failed: " + r->msg);
  return Y3G_OK;
}

int y3g_server_stats_get(y3g_server* s, y3g_server_stats* st)
{
  if (!s || !st) return srv_fail(Y3G_ERR_ARG, "y3g_server_stats_get: bad arguments");
  std::lock_guard<std::mutex> lk(s->mu);
  *st = s->stats;
  st->queued = (long long)s->queue.size();
  st->launches = y3g::stream_launches(s->st);
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
  s->poller.join();  // the queued and in-flight requests finish first
  long long nev = 0;
  int const rc = y3g::stream_close(s->st, &nev);
  y3g_runtime_destroy(s->rt);
  delete s;
  return rc == Y3G_OK ? Y3G_OK : srv_fail(rc, y3g_last_error());
}

}  // extern "C"
