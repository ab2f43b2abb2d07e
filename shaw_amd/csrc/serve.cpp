// Continuous-batching ROM server on one GPU (host C++).
//
// Requests carry m forcing realizations (columns), a ROM size (an engine), a
// horizon and a seismogram sampling period, and arrive at given times. Each
// engine is a resident reduced model (operators, receiver rows and column
// slots with their states, forcing series and seismograms in HBM; a shaw_rom
// handle on the GPU). The loop
//   1. queues the requests that have arrived,
//   2. admits queued requests while their engine has free slots and the
//      resident footprint stays within the HBM budget; requests admitted
//      while others are mid-flight join the running batch (continuous
//      batching). FIFO admits in arrival order and does not pass a request
//      that does not fit. LEAST_WORK admits, among the requests that fit,
//      the one with the least work (horizon x columns) first; a request
//      that has waited max_wait seconds is admitted before any other, and
//      no request for its engine passes it while it does not fit (none at
//      all when the HBM budget is what it waits for), so no request waits
//      without bound under sustained load, and the other engines keep
//      admitting,
//   3. advances every idle engine with active columns by one chunk of steps
//      (each engine on its own stream; under a time budget, by the time of
//      that many one-steps, see shaw_rom_set_time_budget), so the engines'
//      work overlaps on the device, and
//   4. when an engine's chunk is done (a non-blocking query), counts its
//      columns' steps (the engine's steps call, when it has one) and
//      retires its requests that reached their horizon (latency =
//      completion - arrival); the engine is then admitted into and advanced
//      again without waiting for the others.
// With one step at a time a column's arithmetic does not depend on the
// others in the batch, so a request's result is the same alone and in the
// mix (with macro steps, to the BF16 rounding).
// The reference has no serving loop; its ROM entry point runs one problem
// per process (src/kokkos/main.cc:66-73, rom_problem_rank_one.hpp:123-159).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "shaw_common.hpp"

namespace shaw {
namespace {

struct Live {
  uint64_t idx;        // request index
  uint64_t left;       // steps to go
  uint64_t bytes;      // footprint
  std::vector<uint64_t> slots;
};

uint64_t columnBytes(const shaw_engine& e, uint64_t horizon, uint64_t every) {
  const uint64_t samples = every ? horizon / every : 0;
  return 8 * (e.kv + e.ks + e.kv + horizon + e.r * samples);
}

// numpy's default percentile: linear interpolation between order statistics
double percentile(std::vector<double> x, double p) {
  if (x.empty()) return 0.0;
  std::sort(x.begin(), x.end());
  const double pos = p / 100.0 * (double)(x.size() - 1);
  const size_t lo = (size_t)std::floor(pos), hi = std::min(lo + 1, x.size() - 1);
  return x[lo] + (x[hi] - x[lo]) * (pos - (double)lo);
}

void check(int rc, const char* what) {
  SHAW_REQUIRE(rc == SHAW_OK, rc, std::string(what) + " failed: " + shaw_last_error());
}

}  // namespace

void serve(const shaw_engine* eng, int ne, const shaw_serve_options& o, shaw_request* req, uint64_t n,
           shaw_serve_stats* st) {
  SHAW_REQUIRE(eng && ne > 0 && (req || n == 0), SHAW_ERR_ARG, "bad argument");
  SHAW_REQUIRE(o.chunk_steps > 0, SHAW_ERR_ARG, "chunk_steps must be > 0");
  SHAW_REQUIRE(o.policy == SHAW_SERVE_FIFO || o.policy == SHAW_SERVE_LEAST_WORK, SHAW_ERR_ARG,
               "unknown admission policy");
  const auto t0 = std::chrono::steady_clock::now();
  auto now = [&]() -> double {
    if (o.clock) return o.clock(o.clock_ctx);
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  };
  const double start = now();
  auto rel = [&]() { return now() - start; };
  std::vector<uint64_t> footprint(n);
  for (uint64_t i = 0; i < n; ++i) {
    shaw_request& r = req[i];
    SHAW_REQUIRE(r.engine >= 0 && r.engine < ne, SHAW_ERR_ARG, "request " + std::to_string(r.id) + ": no such engine");
    const shaw_engine& e = eng[r.engine];
    SHAW_REQUIRE(r.m > 0 && r.m <= e.capacity, SHAW_ERR_ARG,
                 "request " + std::to_string(r.id) + ": columns exceed the engine capacity");
    SHAW_REQUIRE(r.horizon > 0 && r.b && (r.ftab || (r.kinds && r.delays && r.periods)), SHAW_ERR_ARG,
                 "request " + std::to_string(r.id) + ": no inputs");
    SHAW_REQUIRE(r.ftab || e.admit_signals, SHAW_ERR_ARG,
                 "request " + std::to_string(r.id) + ": its engine takes forcing tables only");
    footprint[i] = r.m * columnBytes(e, r.horizon, r.every);
    SHAW_REQUIRE(footprint[i] <= o.hbm_budget, SHAW_ERR_ARG,
                 "request " + std::to_string(r.id) + ": footprint exceeds the HBM budget");
    r.admitted_s = r.done_s = -1.0;
  }
  std::vector<uint64_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return req[a].arrival_s < req[b].arrival_s; });

  std::deque<uint64_t> queue;
  std::vector<Live> live;
  std::vector<uint64_t> free(ne), launched(ne, 0);
  std::vector<char> busy(ne, 0);
  uint64_t resident = 0, forcingSteps = 0, launches = 0, maxQueue = 0, finished = 0;
  double maxWait = 0.0;
  auto refreshFree = [&](int k) {
    uint64_t f = 0;
    check(eng[k].free_slots(eng[k].ctx, &f), "free_slots");
    free[k] = f;
  };
  for (int k = 0; k < ne; ++k) refreshFree(k);
  // admission needs the request's engine idle (an admission waits for the
  // engine's launches in flight)
  auto fits = [&](uint64_t i) {
    const int k = req[i].engine;
    return !busy[k] && free[k] >= req[i].m && resident + footprint[i] <= o.hbm_budget;
  };
  auto admitOne = [&](std::deque<uint64_t>::iterator it) {
    const uint64_t i = *it;
    queue.erase(it);
    shaw_request& r = req[i];
    const shaw_engine& e = eng[r.engine];
    Live L{i, r.horizon, footprint[i], std::vector<uint64_t>(r.m)};
    std::vector<uint64_t> hz(r.m, r.horizon);
    if (r.ftab)
      check(e.admit(e.ctx, r.m, r.b, r.horizon, r.ftab, hz.data(), r.every, L.slots.data()), "admit");
    else
      check(e.admit_signals(e.ctx, r.m, r.b, r.kinds, r.delays, r.periods, r.horizon, hz.data(), r.every,
                            L.slots.data()),
            "admit");
    refreshFree(r.engine);
    r.admitted_s = rel();
    maxWait = std::max(maxWait, r.admitted_s - r.arrival_s);
    resident += L.bytes;
    live.push_back(std::move(L));
  };
  std::vector<char> held(ne);
  auto admit = [&]() {
    while (!queue.empty()) {
      if (o.policy == SHAW_SERVE_FIFO) {
        // in arrival order; the head is not passed
        if (!fits(queue.front())) return;
        admitOne(queue.begin());
        continue;
      }
      // least work first, but requests past their wait bound go first, in
      // arrival order; one that does not fit holds its engine (no request
      // passes it there), or every engine when the HBM budget is short
      std::fill(held.begin(), held.end(), 0);
      bool admitted = false;
      if (o.max_wait_s > 0) {
        for (auto it = queue.begin(); it != queue.end(); ++it) {  // the queue is in arrival order
          const uint64_t i = *it;
          if (rel() - req[i].arrival_s < o.max_wait_s) break;
          if (held[req[i].engine]) continue;
          if (fits(i)) {
            admitOne(it);
            admitted = true;
            break;
          }
          if (resident + footprint[i] > o.hbm_budget) return;
          held[req[i].engine] = 1;
        }
      }
      if (admitted) continue;
      auto best = queue.end();
      for (auto it = queue.begin(); it != queue.end(); ++it) {
        if (held[req[*it].engine] || !fits(*it)) continue;
        if (best == queue.end() || req[*it].horizon * req[*it].m < req[*best].horizon * req[*best].m) best = it;
      }
      if (best == queue.end()) return;
      admitOne(best);
    }
  };
  // an engine whose chunk finished: count the steps its columns took (the
  // steps asked for, or under a time budget what the engine reports: a
  // request is as far as its slowest column), retire the requests that
  // reached their horizon
  std::vector<uint64_t> slotsOf, stepsOf;
  auto harvest = [&](int k) {
    // the step counts of all the engine's live columns, in one call
    if (eng[k].steps) {
      slotsOf.clear();
      for (const Live& L : live)
        if (req[L.idx].engine == k) slotsOf.insert(slotsOf.end(), L.slots.begin(), L.slots.end());
      stepsOf.resize(slotsOf.size());
      if (!slotsOf.empty()) check(eng[k].steps(eng[k].ctx, slotsOf.size(), slotsOf.data(), stepsOf.data()), "steps");
    }
    std::vector<Live> keep;
    size_t at = 0;  // this request's columns in stepsOf
    for (Live& L : live) {
      shaw_request& r = req[L.idx];
      if (r.engine != k) {
        keep.push_back(std::move(L));
        continue;
      }
      uint64_t taken = std::min(launched[k], L.left);
      if (eng[k].steps) {
        const uint64_t done = std::min(*std::min_element(stepsOf.begin() + at, stepsOf.begin() + at + r.m), r.horizon);
        at += r.m;
        const uint64_t before = r.horizon - L.left;
        taken = done > before ? done - before : 0;
      }
      L.left -= taken;
      forcingSteps += taken * r.m;
      if (L.left > 0) {
        keep.push_back(std::move(L));
        continue;
      }
      const shaw_engine& e = eng[k];
      const uint64_t T = r.every ? r.horizon / r.every : 0;
      // seis NULL: the seismograms are not wanted and are not copied out
      check(e.retire_batch(e.ctx, r.m, L.slots.data(), T, r.seis, nullptr), "retire");
      r.done_s = rel();
      resident -= L.bytes;
      ++finished;
    }
    live = std::move(keep);
    launched[k] = 0;
    refreshFree(k);
  };
  uint64_t next = 0;
  while (finished < n) {
    const double t = rel();
    while (next < n && req[order[next]].arrival_s <= t) queue.push_back(order[next++]);
    maxQueue = std::max<uint64_t>(maxQueue, queue.size());
    bool progressed = false;
    for (int k = 0; k < ne; ++k) {
      if (!busy[k]) continue;
      int idle = 0;
      check(eng[k].query ? eng[k].query(eng[k].ctx, &idle) : eng[k].sync(eng[k].ctx), "query");
      if (!eng[k].query) idle = 1;
      if (!idle) continue;
      busy[k] = 0;
      harvest(k);
      progressed = true;
    }
    if (queue.empty() && live.empty() && finished < n) {
      // idle until the next arrival
      const double wait = req[order[next]].arrival_s - t;
      if (o.wait_until) o.wait_until(o.clock_ctx, start + req[order[next]].arrival_s);
      else std::this_thread::sleep_for(std::chrono::duration<double>(std::min(0.01, std::max(0.0, wait))));
      continue;
    }
    admit();
    // every idle engine with active columns advances by one chunk
    for (int k = 0; k < ne; ++k) {
      if (busy[k]) continue;
      uint64_t left = 0;
      for (const Live& L : live)
        if (req[L.idx].engine == k) left = std::max(left, L.left);
      if (!left) continue;
      launched[k] = std::min<uint64_t>(left, o.chunk_steps);
      check(eng[k].advance(eng[k].ctx, launched[k]), "advance");
      busy[k] = 1;
      ++launches;
      progressed = true;
    }
    if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
    SHAW_REQUIRE(o.wall_limit_s <= 0 || rel() <= o.wall_limit_s, SHAW_ERR_STATE,
                 "serving run exceeded its wall-clock limit");
  }
  if (st) {
    std::vector<double> lat;
    double first = 1e300, last = -1e300;
    for (uint64_t i = 0; i < n; ++i) {
      lat.push_back(req[i].done_s - req[i].arrival_s);
      first = std::min(first, req[i].arrival_s);
      last = std::max(last, req[i].done_s);
    }
    st->requests = n;
    st->forcing_steps = forcingSteps;
    st->launches = launches;
    st->max_queue = maxQueue;
    st->span_s = n ? last - first : 0.0;
    st->p50_latency_s = percentile(lat, 50.0);
    st->p99_latency_s = percentile(lat, 99.0);
    st->max_latency_s = lat.empty() ? 0.0 : *std::max_element(lat.begin(), lat.end());
    st->max_wait_s = maxWait;
    st->throughput = st->span_s > 0 ? (double)forcingSteps / st->span_s : 0.0;
  }
}

}  // namespace shaw

namespace {
int romAdmit(void* c, uint64_t n, const double* b, uint64_t nsteps, const double* ftab, const uint64_t* h,
             uint64_t every, uint64_t* slots) {
  return shaw_rom_admit(static_cast<shaw_rom_t>(c), n, b, nsteps, ftab, h, every, slots);
}
int romAdmitSignals(void* c, uint64_t n, const double* b, const int* kinds, const double* delays,
                    const double* periods, uint64_t nsteps, const uint64_t* h, uint64_t every, uint64_t* slots) {
  return shaw_rom_admit_signals(static_cast<shaw_rom_t>(c), n, b, kinds, delays, periods, nsteps, h, every, slots);
}
int romAdvance(void* c, uint64_t n) { return shaw_rom_advance(static_cast<shaw_rom_t>(c), n); }
int romSync(void* c) { return shaw_rom_sync(static_cast<shaw_rom_t>(c)); }
int romRetire(void* c, uint64_t n, const uint64_t* slots, uint64_t maxSamples, double* seis, uint64_t* samples) {
  return shaw_rom_retire_batch(static_cast<shaw_rom_t>(c), n, slots, maxSamples, seis, samples);
}
int romFree(void* c, uint64_t* n) { return shaw_rom_free_slots(static_cast<shaw_rom_t>(c), n); }
int romQuery(void* c, int* idle) { return shaw_rom_query(static_cast<shaw_rom_t>(c), idle); }
int romSteps(void* c, uint64_t n, const uint64_t* slots, uint64_t* steps) {
  return shaw_rom_steps(static_cast<shaw_rom_t>(c), n, slots, steps);
}
}  // namespace

extern "C" {

int shaw_engine_from_rom(shaw_rom_t rom, shaw_engine* out) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(rom && out, SHAW_ERR_ARG, "null argument");
    shaw_engine e{};
    e.ctx = rom;
    uint64_t maxSamples = 0;
    int rc = shaw_rom_info(rom, &e.kv, &e.ks, &e.r, &e.capacity, &maxSamples);
    SHAW_REQUIRE(rc == SHAW_OK, rc, shaw_last_error());
    e.admit = romAdmit;
    e.advance = romAdvance;
    e.sync = romSync;
    e.retire_batch = romRetire;
    e.free_slots = romFree;
    e.query = romQuery;
    e.admit_signals = romAdmitSignals;
    e.steps = romSteps;
    *out = e;
  });
}

int shaw_serve(const shaw_engine* engines, int n_engines, const shaw_serve_options* opts, shaw_request* reqs,
               uint64_t n, shaw_serve_stats* stats) {
  using namespace shaw;
  return guarded([&] {
    SHAW_REQUIRE(opts, SHAW_ERR_ARG, "null options");
    serve(engines, n_engines, *opts, reqs, n, stats);
  });
}

}  // extern "C"
