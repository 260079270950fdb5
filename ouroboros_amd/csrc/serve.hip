// The serving form: a host request loop in front of one device-resident pool. Requests arrive as a
// Poisson process; each is one allocation-heavy kernel (k_request). Requests that arrive while
// others run are admitted as soon as a slot (a stream) is free and the pool's free-chunk count, read
// from the device, covers their estimated demand, so their kernels overlap on the device. Latency is
// arrival to kernel completion on the device clock.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <thread>
#include <vector>

#include "manager.hpp"

using namespace ouro_lib;

#define OURO_TRY_HIP(x)                      \
  do {                                       \
    const hipError_t _e = (x);               \
    if (_e != hipSuccess) return OURO_E_HIP; \
  } while (0)

namespace {

struct Request {
  double arrival_ms;  // since the start of the run
  uint32_t threads;
  uint64_t seed;
  int slot = -1;
  double done_ms = 0;
};

struct Guard {
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;
  ~Guard() {
    for (auto e : events) (void)hipEventDestroy(e);
    for (auto s : streams) (void)hipStreamDestroy(s);
  }
};

struct Pinned {
  void* p = nullptr;
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
};

}  // namespace

extern "C" {

int ouro_serve(ouro_handle h, const ouro_serve_config* cfg, ouro_serve_result* out) {
  if (!h || !cfg || !out || cfg->requests == 0 || cfg->rate <= 0 || cfg->slots == 0 || cfg->slots > 64 ||
      cfg->min_log2 > cfg->max_log2 || cfg->max_log2 > 24 || cfg->classes == 0 || cfg->classes > 10)
    return OURO_E_INVALID;
  if (cfg->agg != 1 && cfg->agg != 16 && cfg->agg != 64) return OURO_E_INVALID;
  memset(out, 0, sizeof(*out));
  IManager& m = *h->impl;
  const uint32_t R = cfg->requests, S = cfg->slots;

  // Zipf class table (P(q) proportional to (q + 1)^-s) as 32-bit thresholds
  ZipfTable z{};
  z.classes = cfg->classes;
  std::vector<double> pq(cfg->classes);
  {
    double acc = 0;
    for (uint32_t q = 0; q < cfg->classes; ++q) acc += (pq[q] = std::pow((double)(q + 1), -cfg->zipf));
    double c = 0;
    for (uint32_t q = 0; q < cfg->classes; ++q) {
      pq[q] /= acc;
      c += pq[q];
      z.thr[q] = q + 1 == cfg->classes ? 0xFFFFFFFFu : (uint32_t)std::min(4294967295.0, std::floor(c * 4294967296.0));
    }
  }
  // the arrival schedule: exponential gaps, log-uniform thread counts, a seed per request
  std::vector<Request> req(R);
  double t = 0;
  for (uint32_t r = 0; r < R; ++r) {
    t += -std::log(1.0 - unit(keyed(cfg->seed ^ 0xA0761D6478BD642Full, r))) / cfg->rate * 1e3;
    const double e = cfg->min_log2 + (cfg->max_log2 - cfg->min_log2) * unit(keyed(cfg->seed ^ 0xE7037ED1A0B428DBull, r));
    req[r].arrival_ms = t;
    req[r].threads = (uint32_t)std::floor(std::exp2(e));
    req[r].seed = keyed(cfg->seed ^ 0x8EBC6AF09C88C6E3ull, r);
  }

  // pool classes and their free pages, from a copy of the device state
  const size_t state_bytes = m.state_bytes();
  Pinned probe_buf;
  OURO_TRY_HIP(hipHostMalloc(&probe_buf.p, state_bytes, hipHostMallocDefault));
  const void* d_state = m.device_manager();
  OURO_TRY_HIP(hipDeviceSynchronize());
  OURO_TRY_HIP(hipMemcpy(probe_buf.p, d_state, state_bytes, hipMemcpyDeviceToHost));
  PoolFree pf{};
  m.pool_free(probe_buf.p, &pf);
  // request class q (16 B << q) lands in the first pool class whose pages hold it
  std::vector<uint32_t> cls(cfg->classes);
  for (uint32_t q = 0; q < cfg->classes; ++q) {
    uint32_t k = 0;
    while (k < pf.classes && pf.page_size[k] < (16u << q)) ++k;
    if (k == pf.classes) return OURO_E_INVALID;
    cls[q] = k;
  }
  // page demand per pool class: the expected count plus four standard deviations, so that the
  // admission test holds with overwhelming probability whatever the draws
  const uint32_t NK = pf.classes;
  std::vector<double> need((size_t)R * NK, 0.0);
  for (uint32_t r = 0; r < R; ++r)
    for (uint32_t q = 0; q < cfg->classes; ++q) {
      const double mean = req[r].threads * pq[q];
      need[(size_t)r * NK + cls[q]] += std::ceil(mean + 4.0 * std::sqrt(mean) + 1.0);
    }

  Guard gd;
  gd.streams.resize(S + 1);
  for (auto& s : gd.streams) OURO_TRY_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  const hipStream_t probe_stream = gd.streams[S];
  // per request: a completion event and a result line of 4 counters
  hipEvent_t base, probe_ev;
  OURO_TRY_HIP(hipEventCreate(&base));
  OURO_TRY_HIP(hipEventCreate(&probe_ev));
  gd.events.push_back(base);
  gd.events.push_back(probe_ev);
  std::vector<hipEvent_t> req_done(R);
  for (auto& e : req_done) OURO_TRY_HIP(hipEventCreate(&e));
  gd.events.insert(gd.events.end(), req_done.begin(), req_done.end());
  unsigned long long* d_res = nullptr;
  OURO_TRY_HIP(hipMalloc(&d_res, (size_t)R * 4 * sizeof(unsigned long long)));
  std::unique_ptr<void, void (*)(void*)> res_guard(d_res, [](void* p) { (void)hipFree(p); });
  OURO_TRY_HIP(hipMemset(d_res, 0, (size_t)R * 4 * sizeof(unsigned long long)));
  OURO_TRY_HIP(hipDeviceSynchronize());

  // the device-resident free capacity, refreshed by a copy of the manager state in flight
  bool probe_busy = false, stale = false;
  auto issue_probe = [&]() -> hipError_t {
    hipError_t e = hipMemcpyAsync(probe_buf.p, d_state, state_bytes, hipMemcpyDeviceToHost, probe_stream);
    if (e == hipSuccess) e = hipEventRecord(probe_ev, probe_stream);
    probe_busy = e == hipSuccess;
    return e;
  };
  // Admission: the snapshot's free pages must cover the full demand of every request in flight
  // plus the new one, class by class, with the shortfall taken in whole chunks from the shared
  // region. A snapshot taken while a request runs already lacks what that request holds, and its
  // full demand is counted again, so the test errs on the safe side whenever the snapshot was taken.
  std::vector<int> slot_req(S, -1);
  std::vector<double> total(NK);
  auto fits = [&](uint32_t r) {
    for (uint32_t k = 0; k < NK; ++k) total[k] = need[(size_t)r * NK + k];
    for (uint32_t s = 0; s < S; ++s)
      if (slot_req[s] >= 0)
        for (uint32_t k = 0; k < NK; ++k) total[k] += need[(size_t)slot_req[s] * NK + k];
    double chunks = 0;
    for (uint32_t k = 0; k < NK; ++k) {
      const double short_pages = total[k] - (double)std::max<int64_t>(0, pf.pages[k]);
      if (short_pages > 0) chunks += std::ceil(std::ceil(short_pages / pf.ppc[k]) * pf.cost[k]);
    }
    return chunks <= (double)pf.shared;
  };

  OURO_TRY_HIP(hipEventRecord(base, gd.streams[0]));
  OURO_TRY_HIP(hipEventSynchronize(base));
  const auto t0 = std::chrono::steady_clock::now();
  auto now_ms = [&]() { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); };
  uint32_t next = 0, completed = 0, inflight = 0;
  std::vector<uint8_t> deferred(R, 0);
  while (completed < R) {
    bool progress = false;
    // completions
    for (uint32_t s = 0; s < S; ++s) {
      if (slot_req[s] < 0) continue;
      const hipError_t q = hipEventQuery(req_done[slot_req[s]]);
      if (q == hipErrorNotReady) continue;
      OURO_TRY_HIP(q);
      slot_req[s] = -1;
      --inflight;
      ++completed;
      progress = true;
      stale = true;
    }
    // a fresh snapshot once requests have finished (their frees are not in the last one)
    if (probe_busy) {
      const hipError_t q = hipEventQuery(probe_ev);
      if (q == hipSuccess) {
        m.pool_free(probe_buf.p, &pf);
        probe_busy = false;
      } else if (q != hipErrorNotReady) {
        return OURO_E_HIP;
      }
    }
    if (!probe_busy && stale && cfg->admit) {
      OURO_TRY_HIP(issue_probe());
      stale = false;
    }
    // admissions, in arrival order
    const double now = now_ms();
    while (next < R && req[next].arrival_ms <= now && inflight < S) {
      if (cfg->admit && inflight > 0 && !fits(next)) {
        if (!deferred[next]) ++out->deferred;
        deferred[next] = 1;
        break;
      }
      uint32_t s = 0;
      while (slot_req[s] >= 0) ++s;
      Request& rq = req[next];
      OURO_TRY_HIP(m.launch_request(rq.seed, rq.threads, z, d_res + (size_t)next * 4, cfg->agg, gd.streams[s]));
      OURO_TRY_HIP(hipEventRecord(req_done[next], gd.streams[s]));
      rq.slot = (int)s;
      slot_req[s] = (int)next;
      ++inflight;
      out->max_in_flight = std::max(out->max_in_flight, inflight);
      ++next;
      progress = true;
    }
    if (!progress) {
      // nothing finished or started: wait for the earlier of the next arrival and a completion
      if (inflight == 0 && next < R) {
        const double wait = req[next].arrival_ms - now_ms();
        if (wait > 0.2) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)((wait - 0.1) * 1e3)));
      } else {
        std::this_thread::yield();
      }
    }
  }
  OURO_TRY_HIP(hipDeviceSynchronize());

  // latencies on the device clock: completion - arrival
  std::vector<double> lat(R);
  double last = 0;
  for (uint32_t r = 0; r < R; ++r) {
    float ms = 0;
    OURO_TRY_HIP(hipEventElapsedTime(&ms, base, req_done[r]));
    req[r].done_ms = ms;
    lat[r] = std::max(0.0, (double)ms - req[r].arrival_ms);
    last = std::max(last, (double)ms);
  }
  std::vector<unsigned long long> res((size_t)R * 4);
  OURO_TRY_HIP(hipMemcpy(res.data(), d_res, res.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  for (uint32_t r = 0; r < R; ++r) {
    res[(size_t)r * 4 + 3] = req[r].threads;
    out->mallocs += req[r].threads;
    out->mismatches += res[(size_t)r * 4 + 1];
    out->nulls += res[(size_t)r * 4 + 2];
    if (cfg->latency_out) cfg->latency_out[r] = lat[r];
  }
  if (cfg->results_out) memcpy(cfg->results_out, res.data(), res.size() * sizeof(unsigned long long));
  std::vector<double> sorted = lat;
  std::sort(sorted.begin(), sorted.end());
  auto pct = [&](double p) { return sorted[(size_t)std::min<double>(R - 1, std::floor(p * (R - 1) + 0.5))]; };
  out->p50_ms = pct(0.50);
  out->p99_ms = pct(0.99);
  out->max_ms = sorted.back();
  double sum = 0;
  for (double v : lat) sum += v;
  out->mean_ms = sum / R;
  out->wall_ms = last - req[0].arrival_ms;
  out->mops = out->wall_ms > 0 ? 2.0 * (double)out->mallocs / (out->wall_ms * 1e-3) / 1e6 : 0.0;
  out->achieved_rate = out->wall_ms > 0 ? R / (out->wall_ms * 1e-3) : 0.0;
  out->offered_ms = req[R - 1].arrival_ms - req[0].arrival_ms;
  out->completed = completed;
  ouro_state st;
  if (m.query(&st) != hipSuccess) return OURO_E_HIP;
  out->error_word = st.error;
  return OURO_OK;
}

int ouro_serve_expected(const ouro_serve_config* cfg, uint32_t r, uint64_t* threads, uint64_t* words) {
  // words a request reads back intact when nothing interferes: the sum of its blocks' words
  if (!cfg || !threads || !words || cfg->classes == 0 || cfg->classes > 10) return OURO_E_INVALID;
  std::vector<double> pq(cfg->classes);
  uint32_t thr[16];
  double acc = 0;
  for (uint32_t q = 0; q < cfg->classes; ++q) acc += (pq[q] = std::pow((double)(q + 1), -cfg->zipf));
  double c = 0;
  for (uint32_t q = 0; q < cfg->classes; ++q) {
    c += pq[q] / acc;
    thr[q] = q + 1 == cfg->classes ? 0xFFFFFFFFu : (uint32_t)std::min(4294967295.0, std::floor(c * 4294967296.0));
  }
  const double e = cfg->min_log2 + (cfg->max_log2 - cfg->min_log2) * unit(keyed(cfg->seed ^ 0xE7037ED1A0B428DBull, r));
  const uint32_t n = (uint32_t)std::floor(std::exp2(e));
  const uint64_t seed = keyed(cfg->seed ^ 0x8EBC6AF09C88C6E3ull, r);
  uint64_t w = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t x = (uint32_t)(keyed(seed, i) >> 32);
    uint32_t q = 0;
    while (q + 1 < cfg->classes && x >= thr[q]) ++q;
    w += (16u << q) / 4u;
  }
  *threads = n;
  *words = w;
  return OURO_OK;
}

}  // extern "C"
