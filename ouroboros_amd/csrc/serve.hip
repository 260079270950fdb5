// The serving form: a host request loop in front of one device-resident pool, with continuous
// batching. Requests arrive as a Poisson process; each is an allocation-heavy piece of work (every
// thread allocates, writes, verifies and frees a block). Whenever a launch slot (a stream) is free,
// the requests that have arrived and fit the pool are gathered into one launch — each workgroup of
// it serves a stretch of one request — so a request arriving while others run joins the next launch
// instead of waiting for them, and launches overlap on the device. Admission uses the pool's free
// pages, read from the device-resident manager state by an asynchronous copy. Latency is arrival to
// the completion of the request's launch, on the device clock.
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
  int batch = -1;     // launch that served it
};

struct Batch {
  uint32_t lo, hi;    // requests lo .. hi (launches take requests in arrival order)
  hipEvent_t done;
};

struct Guard {
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;
  std::vector<void*> host, dev;
  ~Guard() {
    for (auto e : events) (void)hipEventDestroy(e);
    for (auto s : streams) (void)hipStreamDestroy(s);
    for (auto p : host) (void)hipHostFree(p);
    for (auto p : dev) (void)hipFree(p);
  }
};

// Zipf class probabilities P(q) proportional to (q + 1)^-s and the device's 32-bit thresholds.
void zipf_table(const ouro_serve_config* cfg, std::vector<double>& pq, ZipfTable& z) {
  z = ZipfTable{};
  z.classes = cfg->classes;
  pq.assign(cfg->classes, 0.0);
  double acc = 0;
  for (uint32_t q = 0; q < cfg->classes; ++q) acc += (pq[q] = std::pow((double)(q + 1), -cfg->zipf));
  double c = 0;
  for (uint32_t q = 0; q < cfg->classes; ++q) {
    pq[q] /= acc;
    c += pq[q];
    z.thr[q] = q + 1 == cfg->classes ? 0xFFFFFFFFu : (uint32_t)std::min(4294967295.0, std::floor(c * 4294967296.0));
  }
}

uint32_t request_threads(const ouro_serve_config* cfg, uint32_t r) {
  const double e = cfg->min_log2 + (cfg->max_log2 - cfg->min_log2) * unit(keyed(cfg->seed ^ 0xE7037ED1A0B428DBull, r));
  return (uint32_t)std::floor(std::exp2(e));
}

uint64_t request_seed(const ouro_serve_config* cfg, uint32_t r) { return keyed(cfg->seed ^ 0x8EBC6AF09C88C6E3ull, r); }

}  // namespace

extern "C" {

int ouro_serve(ouro_handle h, const ouro_serve_config* cfg, ouro_serve_result* out) {
  if (!h || !cfg || !out || cfg->requests == 0 || cfg->rate <= 0 || cfg->slots == 0 || cfg->slots > 64 ||
      cfg->min_log2 > cfg->max_log2 || cfg->max_log2 > 24 || cfg->classes == 0 || cfg->classes > 10)
    return OURO_E_INVALID;
  if (cfg->agg != 0 && cfg->agg != 1 && cfg->agg != 16 && cfg->agg != 64) return OURO_E_INVALID;
  const uint32_t tile = cfg->agg == 0 && cfg->tile ? cfg->tile : 1u;
  if (tile != 1 && tile != 2 && tile != 4) return OURO_E_INVALID;
  memset(out, 0, sizeof(*out));
  IManager& m = *h->impl;
  const uint32_t R = cfg->requests, S = cfg->slots;
  const uint32_t max_batch = cfg->max_batch ? cfg->max_batch : 1u;
  // threads per workgroup: 256 also for block-wide allocation, where smaller workgroups keep more
  // of them resident per CU to overlap the requests' memory traffic (measured: 1024-thread
  // workgroups served 26% fewer requests per second at saturation); B = request threads per
  // workgroup, `tile` per thread when block-wide
  const uint32_t B = 256u * tile;

  ZipfTable z;
  std::vector<double> pq;
  zipf_table(cfg, pq, z);
  // the arrival schedule: exponential gaps, log-uniform thread counts, a seed per request
  std::vector<Request> req(R);
  double t = 0;
  uint32_t most_blocks = 0;
  for (uint32_t r = 0; r < R; ++r) {
    t += -std::log(1.0 - unit(keyed(cfg->seed ^ 0xA0761D6478BD642Full, r))) / cfg->rate * 1e3;
    req[r].arrival_ms = t;
    req[r].threads = request_threads(cfg, r);
    req[r].seed = request_seed(cfg, r);
    most_blocks = std::max(most_blocks, (req[r].threads + B - 1) / B);
  }
  // workgroups one launch may carry: about 4M threads, or one whole request if that is larger
  const uint32_t cap_blocks = std::max<uint32_t>((1u << 22) / B, most_blocks);

  // pool classes and their free pages, from a copy of the device state
  const size_t state_bytes = m.state_bytes();
  Guard gd;
  void* probe_buf = nullptr;
  OURO_TRY_HIP(hipHostMalloc(&probe_buf, state_bytes, hipHostMallocDefault));
  gd.host.push_back(probe_buf);
  const void* d_state = m.device_manager();
  OURO_TRY_HIP(hipDeviceSynchronize());
  OURO_TRY_HIP(hipMemcpy(probe_buf, d_state, state_bytes, hipMemcpyDeviceToHost));
  PoolFree pf{};
  m.pool_free(probe_buf, &pf);
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

  gd.streams.resize(S + 1);
  for (auto& st : gd.streams) OURO_TRY_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  const hipStream_t probe_stream = gd.streams[S];
  hipEvent_t base, probe_ev;
  OURO_TRY_HIP(hipEventCreate(&base));
  gd.events.push_back(base);
  OURO_TRY_HIP(hipEventCreate(&probe_ev));
  gd.events.push_back(probe_ev);
  // per slot: pinned descriptor staging and its device copy
  std::vector<RequestBlock*> h_desc(S), d_desc(S);
  for (uint32_t i = 0; i < S; ++i) {
    void* hp = nullptr;
    void* dp = nullptr;
    OURO_TRY_HIP(hipHostMalloc(&hp, (size_t)cap_blocks * sizeof(RequestBlock), hipHostMallocDefault));
    gd.host.push_back(hp);
    OURO_TRY_HIP(hipMalloc(&dp, (size_t)cap_blocks * sizeof(RequestBlock)));
    gd.dev.push_back(dp);
    h_desc[i] = static_cast<RequestBlock*>(hp);
    d_desc[i] = static_cast<RequestBlock*>(dp);
  }
  unsigned long long* d_res = nullptr;
  OURO_TRY_HIP(hipMalloc(&d_res, (size_t)R * 4 * sizeof(unsigned long long)));
  gd.dev.push_back(d_res);
  OURO_TRY_HIP(hipMemset(d_res, 0, (size_t)R * 4 * sizeof(unsigned long long)));
  OURO_TRY_HIP(hipDeviceSynchronize());

  // the device-resident free capacity, refreshed by a copy of the manager state in flight
  bool probe_busy = false, stale = false;
  auto issue_probe = [&]() -> hipError_t {
    hipError_t e = hipMemcpyAsync(probe_buf, d_state, state_bytes, hipMemcpyDeviceToHost, probe_stream);
    if (e == hipSuccess) e = hipEventRecord(probe_ev, probe_stream);
    probe_busy = e == hipSuccess;
    return e;
  };
  // Admission: the snapshot's free pages must cover the full demand of every request in flight
  // plus the new one, class by class, with the shortfall taken in whole chunks (with their share of
  // queue chunks) from the shared region. A snapshot taken while a request runs already lacks what
  // that request holds, and its full demand is counted again, so the test errs on the safe side
  // whenever the snapshot was taken.
  std::vector<double> held(NK, 0.0);  // demand of the requests launched and not yet complete
  auto fits = [&](uint32_t r) {
    double chunks = 0;
    for (uint32_t k = 0; k < NK; ++k) {
      const double short_pages = held[k] + need[(size_t)r * NK + k] - (double)std::max<int64_t>(0, pf.pages[k]);
      if (short_pages > 0) chunks += std::ceil(std::ceil(short_pages / pf.ppc[k]) * pf.cost[k]);
    }
    return chunks <= (double)pf.shared;
  };
  auto hold = [&](uint32_t r, double sign) {
    for (uint32_t k = 0; k < NK; ++k) held[k] += sign * need[(size_t)r * NK + k];
  };

  std::vector<Batch> batches;
  batches.reserve(R);
  std::vector<int> slot_batch(S, -1);
  OURO_TRY_HIP(hipEventRecord(base, gd.streams[0]));
  OURO_TRY_HIP(hipEventSynchronize(base));
  const auto t0 = std::chrono::steady_clock::now();
  auto now_ms = [&]() { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); };
  uint32_t next = 0, completed = 0, inflight = 0;
  std::vector<uint8_t> deferred(R, 0);
  while (completed < R) {
    bool progress = false;
    // completions: a launch completes all its requests at once
    for (uint32_t s = 0; s < S; ++s) {
      if (slot_batch[s] < 0) continue;
      const Batch& bt = batches[slot_batch[s]];
      const hipError_t q = hipEventQuery(bt.done);
      if (q == hipErrorNotReady) continue;
      OURO_TRY_HIP(q);
      for (uint32_t r = bt.lo; r < bt.hi; ++r) hold(r, -1.0);
      inflight -= bt.hi - bt.lo;
      completed += bt.hi - bt.lo;
      slot_batch[s] = -1;
      progress = true;
      stale = true;
    }
    // a fresh snapshot once launches have finished (their frees are not in the last one)
    if (probe_busy) {
      const hipError_t q = hipEventQuery(probe_ev);
      if (q == hipSuccess) {
        m.pool_free(probe_buf, &pf);
        probe_busy = false;
      } else if (q != hipErrorNotReady) {
        return OURO_E_HIP;
      }
    }
    if (!probe_busy && stale && cfg->admit) {
      OURO_TRY_HIP(issue_probe());
      stale = false;
    }
    // launches: every free slot takes the arrived requests that fit, in arrival order
    const double now = now_ms();
    for (uint32_t s = 0; s < S && next < R && req[next].arrival_ms <= now; ++s) {
      if (slot_batch[s] >= 0) continue;
      const uint32_t lo = next;
      uint32_t blocks = 0;
      while (next < R && req[next].arrival_ms <= now && next - lo < max_batch) {
        const uint32_t nb = (req[next].threads + B - 1) / B;
        if (blocks + nb > cap_blocks) break;
        if (cfg->admit && (inflight > 0 || next > lo) && !fits(next)) {
          if (!deferred[next]) ++out->deferred;
          deferred[next] = 1;
          break;
        }
        for (uint32_t k = 0; k < nb; ++k) h_desc[s][blocks + k] = RequestBlock{req[next].seed, next, k * B, req[next].threads, 0u};
        blocks += nb;
        hold(next, 1.0);
        req[next].batch = (int)batches.size();
        ++next;
      }
      if (next == lo) break;  // the next request waits for memory or a free slot
      Batch bt{lo, next, nullptr};
      OURO_TRY_HIP(hipEventCreate(&bt.done));
      gd.events.push_back(bt.done);
      OURO_TRY_HIP(hipMemcpyAsync(d_desc[s], h_desc[s], (size_t)blocks * sizeof(RequestBlock), hipMemcpyHostToDevice,
                                  gd.streams[s]));
      OURO_TRY_HIP(m.launch_requests(d_desc[s], blocks, z, d_res, cfg->agg, gd.streams[s], tile));
      OURO_TRY_HIP(hipEventRecord(bt.done, gd.streams[s]));
      slot_batch[s] = (int)batches.size();
      batches.push_back(bt);
      inflight += next - lo;
      out->max_in_flight = std::max(out->max_in_flight, inflight);
      out->launches += 1;
      progress = true;
    }
    if (!progress) {
      // nothing finished or started: sleep until the next arrival if nothing runs, else poll
      if (inflight == 0 && next < R) {
        const double wait = req[next].arrival_ms - now_ms();
        if (wait > 0.2) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)((wait - 0.1) * 1e3)));
      } else {
        std::this_thread::yield();
      }
    }
  }
  OURO_TRY_HIP(hipDeviceSynchronize());

  // latencies on the device clock: completion of the request's launch - arrival
  std::vector<double> lat(R);
  double last = 0;
  std::vector<double> done_ms(batches.size());
  for (size_t b = 0; b < batches.size(); ++b) {
    float ms = 0;
    OURO_TRY_HIP(hipEventElapsedTime(&ms, base, batches[b].done));
    done_ms[b] = ms;
    last = std::max(last, (double)ms);
  }
  for (uint32_t r = 0; r < R; ++r) lat[r] = std::max(0.0, done_ms[req[r].batch] - req[r].arrival_ms);
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
  ZipfTable z;
  std::vector<double> pq;
  zipf_table(cfg, pq, z);
  const uint32_t n = request_threads(cfg, r);
  const uint64_t seed = request_seed(cfg, r);
  uint64_t w = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t x = (uint32_t)(keyed(seed, i) >> 32);
    uint32_t q = 0;
    while (q + 1 < cfg->classes && x >= z.thr[q]) ++q;
    w += (16u << q) / 4u;
  }
  *threads = n;
  *words = w;
  return OURO_OK;
}

}  // extern "C"
