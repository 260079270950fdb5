// extern "C" entry points of libouro.so (declared in include/ouro_capi.h), the trace harness that
// mirrors src/main.cu's alloc / write / verify / free rounds, and the host-side invariant checkers.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <mutex>
#include <new>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "manager.hpp"

using namespace ouro_lib;

IManager* ouro_lib::make_variant(int v) {
  switch (v) {
    case OURO_PQ: return make_pq();
    case OURO_CQ: return make_cq();
    case OURO_VAPQ: return make_vapq();
    case OURO_VACQ: return make_vacq();
    case OURO_VLPQ: return make_vlpq();
    case OURO_VLCQ: return make_vlcq();
    case 6: return make_geo6();
    case 7: return make_geo7();
    case 8: return make_geo8();
    case 9: return make_geo9();
    case 10: return make_geo10();
    case 11: return make_geo11();
    default: return nullptr;
  }
}

namespace {

// Page size a request of `size` bytes receives from a chain whose smallest page is `smallest`.
inline uint32_t page_of(uint32_t size, uint32_t smallest = 16) {
  uint32_t p = smallest;
  while (p < size) p <<= 1;
  return p;
}

// ------------------------------------------------------------------ touch pass (write / verify)
// Writes (or checks) the owning thread's id into every 32-bit word of its block, as
// d_testWriteToMemory / d_testReadFromMemory (src/main.cu:20-53). Blocks up to 256 B are handled by
// their own lane; larger blocks by the whole wavefront with coalesced 16-byte accesses. Pool blocks
// outside [lo, hi) are counted, never dereferenced; blocks of more than `heap_above` bytes come from
// the device heap and need only be 16-byte aligned.
__global__ void __launch_bounds__(256) k_touch(void* const* __restrict__ ptrs, const uint32_t* __restrict__ sizes,
                                               uint32_t n, const uint8_t* lo, const uint8_t* hi, int verify,
                                               unsigned long long* __restrict__ errs, uint32_t heap_above) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = tid < n;
  const uint8_t* p = in ? reinterpret_cast<const uint8_t*>(ptrs[tid]) : nullptr;
  const uint32_t size = in ? sizes[tid] : 0u;
  const uint32_t bytes = size & ~3u;
  const bool ok = p != nullptr && (size > heap_above || (p >= lo && p + bytes <= hi)) &&
                  (reinterpret_cast<uintptr_t>(p) & 15u) == 0;
  if (p != nullptr && !ok) atomicAdd(&errs[1], 1ull);
  if (!verify && in && size && p == nullptr) atomicAdd(&errs[2], 1ull);  // requests answered with nullptr
  unsigned long long bad = 0;
  if (ok && bytes <= 256) {
    uint32_t* w = reinterpret_cast<uint32_t*>(const_cast<uint8_t*>(p));
    for (uint32_t i = 0; i < bytes / 4; ++i) {
      if (!verify) w[i] = tid;
      else bad += (w[i] != tid);
    }
  }
  uint64_t todo = __ballot(ok && bytes > 256);
  const uint32_t ln = __lane_id();
  while (todo) {
    const int l = __ffsll((unsigned long long)todo) - 1;
    todo &= todo - 1;
    const uint64_t addr = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(reinterpret_cast<uint64_t>(p) >> 32), l)
                           << 32) |
                          (uint32_t)__builtin_amdgcn_readlane((int)reinterpret_cast<uint64_t>(p), l);
    const uint32_t nb = (uint32_t)__builtin_amdgcn_readlane((int)bytes, l);
    const uint32_t val = (uint32_t)__builtin_amdgcn_readlane((int)tid, l);
    uint4* b = reinterpret_cast<uint4*>(addr);
    const uint32_t nv = nb / 16;
    for (uint32_t i = ln; i < nv; i += 64) {
      if (!verify) {
        b[i] = make_uint4(val, val, val, val);
      } else {
        const uint4 x = b[i];
        bad += (x.x != val) + (x.y != val) + (x.z != val) + (x.w != val);
      }
    }
  }
  if (verify && bad) atomicAdd(&errs[0], bad);
}

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 4); }
  template <class T>
  T* as() const { return reinterpret_cast<T*>(p); }
};

#define OURO_TRY(x)                          \
  do {                                       \
    hipError_t _e = (x);                     \
    if (_e != hipSuccess) return OURO_E_HIP; \
  } while (0)

// Instances of a variant's chain, from its geometry: the first instance's pages run from the
// smallest page up to its chunk; a chain whose largest page exceeds that chunk has a second instance
// whose pages start at twice the first chunk and whose chunk is its largest page (the MultiOuro
// chains: 16 and 32 KiB pages on 32 KiB chunks of four base chunks, InstanceDefinitions.cuh:17-34).
struct ChainGeo {
  uint32_t smallest[2], largest[2], chunk[2], n;
};
ChainGeo chain_geo(const ouro_variant_info& g) {
  ChainGeo c{};
  c.n = 1;
  c.smallest[0] = g.smallest_page;
  c.largest[0] = c.chunk[0] = g.chunk_bytes;
  if (g.largest_page > g.chunk_bytes) {
    c.n = 2;
    c.smallest[1] = 2 * g.chunk_bytes;
    c.largest[1] = c.chunk[1] = g.largest_page;
  }
  return c;
}

// Run start of a block of page size ps at byte offset `off` from the chunk region, or UINT64_MAX if
// no run of its instance puts a page there: the block must sit on the page grid of a run of
// chunk / 8192 base chunks starting at some base chunk c0 (offset = c0 * 8320 + 128 + p * ps with
// p < chunk / ps), so it never covers the run's header and never reaches past the run's data.
uint64_t run_start_of(uint64_t off, uint32_t ps, uint32_t chunk) {
  const uint64_t idx = off / 8320, span = chunk / 8192;
  for (uint64_t j = 0; j < span && j <= idx; ++j) {
    const uint64_t c0 = idx - j, in = off - c0 * 8320;
    if (in >= 128 && (in - 128) % ps == 0 && (in - 128) + ps <= chunk) return c0;
  }
  return UINT64_MAX;
}

// Host-side check of the live set: every block on the page grid of a run of its instance (any
// geometry), no two live blocks overlapping, and the fragmentation figures (runs holding a live block
// vs. the ideal packing, per class ceil(live blocks / pages per chunk)).
void check_live(const std::vector<uint64_t>& offs, const std::vector<uint32_t>& sizes, ouro_trace_result* r,
                const ouro_variant_info& g) {
  struct Iv {
    uint64_t a, b;
  };
  const ChainGeo cg = chain_geo(g);
  std::vector<Iv> iv;
  iv.reserve(offs.size());
  std::unordered_map<uint32_t, uint64_t> live_ps;  // page size -> live blocks
  std::unordered_set<uint64_t> runs;
  for (size_t i = 0; i < offs.size(); ++i) {
    if (offs[i] == UINT64_MAX || sizes[i] > g.largest_page) continue;  // null, or a device-heap block
    const uint32_t ps = page_of(sizes[i], g.smallest_page);
    iv.push_back({offs[i], offs[i] + ps});
    const uint32_t k = (cg.n == 2 && ps > cg.largest[0]) ? 1u : 0u;
    const uint64_t c0 = ps <= cg.largest[k] ? run_start_of(offs[i], ps, cg.chunk[k]) : UINT64_MAX;
    if (c0 == UINT64_MAX) {
      r->align_errors++;
      continue;
    }
    runs.insert(c0);
    live_ps[ps]++;
  }
  std::sort(iv.begin(), iv.end(), [](const Iv& x, const Iv& y) { return x.a < y.a; });
  for (size_t i = 1; i < iv.size(); ++i)
    if (iv[i].a < iv[i - 1].b) r->overlap_errors++;
  uint64_t ideal = 0;
  for (const auto& [ps, live] : live_ps) {
    const uint32_t k = (cg.n == 2 && ps > cg.largest[0]) ? 1u : 0u;
    const uint64_t ppc = std::max<uint64_t>(1, cg.chunk[k] / ps);
    ideal += (live + ppc - 1) / ppc;
  }
  if (runs.size() > r->peak_live_chunks) {
    r->peak_live_chunks = runs.size();
    r->ideal_chunks = ideal;
  }
}

// ------------------------------------------------------------------ measurement kernels
__global__ void __launch_bounds__(256) k_atomics(uint32_t* buf, uint32_t iters, int pattern, uint32_t* sink) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t* p;
  if (pattern == 0) p = buf + (size_t)tid * 32u;            // one 128-byte line per lane
  else if (pattern == 1) p = buf + (size_t)(tid >> 6) * 32u; // one line per wave
  else p = buf;
  uint32_t acc = 0;
  for (uint32_t i = 0; i < iters; ++i)
    acc += __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (acc == 0xFFFFFFFFu) sink[0] = acc;
}

__global__ void __launch_bounds__(256) k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace

// ------------------------------------------------------------------ allocation-free batch rounds
// Device scratch of ouro_run_batch_static: 256 bytes of error counters, then the request sizes
// (u32) and the block pointers (u64) of up to kStaticRequests requests, 12 MiB in all. Every GPU the
// library runs on holds its own copy until the process ends, and every call on a GPU shares it:
// calls that overlap on one GPU race. Host code never passes `scratch` itself to a kernel or copy
// (in host code the name stands for a host-side shadow, not the device array): it uses the device
// address hipGetSymbolAddress returns, looked up once per device.
constexpr uint32_t kStaticRequests = 1u << 20;
constexpr size_t kStaticCounterBytes = 256;
constexpr size_t kStaticBytes = kStaticCounterBytes + (size_t)kStaticRequests * 12;
static __device__ __align__(256) uint8_t scratch[kStaticBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_dev[kMaxDevices] = {};

// The current device's copy of `scratch` (its device address), looked up on the first call there.
hipError_t static_scratch(uint8_t** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[dev] = reinterpret_cast<uint8_t*>(p);
  }
  *out = g_scratch_dev[dev];
  return hipSuccess;
}

// Clears the counters with a kernel (a fill kernel rather than a memset node, so the launch sequence
// stays valid inside a caller's graph capture).
__global__ void k_zero_counters(unsigned long long* c, uint32_t n) {
  if (threadIdx.x < n) c[threadIdx.x] = 0ull;
}

}  // namespace

extern "C" {

#ifdef OURO_DIAG
int ouro_diag_read(ouro_handle h, unsigned long long* out16, int reset) {
  if (!h || !out16) return OURO_E_INVALID;
  return h->impl->diag_read(out16, reset) == hipSuccess ? OURO_OK : OURO_E_HIP;
}
#endif

int ouro_sem_share(uint64_t before, int64_t num, int64_t ppc, uint32_t* k) {
  if (!k || num < 0 || ppc < 1 || (ppc & (ppc - 1))) return OURO_E_INVALID;
  *k = ouro::sem_share(before, num, ppc);
  return OURO_OK;
}

uint64_t ouro_sem_make(int64_t count, uint64_t expected) { return ouro::Sem::make(count, expected); }

int ouro_abi_sizes(uint32_t* out, uint32_t n) {
  if (!out) return OURO_E_INVALID;
  const uint32_t sizes[] = {sizeof(ouro_variant_info), sizeof(ouro_state),         sizeof(ouro_check_result),
                            sizeof(ouro_trace_config), sizeof(ouro_trace_result),  sizeof(ouro_batch_config),
                            sizeof(ouro_batch_result), sizeof(ouro_batch_workload), sizeof(ouro_serve_config),
                            sizeof(ouro_serve_result)};
  for (uint32_t i = 0; i < n && i < sizeof(sizes) / sizeof(sizes[0]); ++i) out[i] = sizes[i];
  return OURO_OK;
}

int ouro_bench_atomics(uint32_t threads, uint32_t iters, int pattern, double* ops_per_s) {
  if (!ops_per_s || threads == 0 || threads % 256 || iters == 0 || pattern < 0 || pattern > 2) return OURO_E_INVALID;
  DevBuf buf, sink;
  const size_t words = pattern == 0 ? (size_t)threads * 32 : (pattern == 1 ? (size_t)(threads / 64) * 32 : 32);
  OURO_TRY(buf.alloc(words * 4));
  OURO_TRY(sink.alloc(4));
  OURO_TRY(hipMemset(buf.p, 0, words * 4));
  hipEvent_t a, b;
  OURO_TRY(hipEventCreate(&a));
  OURO_TRY(hipEventCreate(&b));
  hipLaunchKernelGGL(k_atomics, dim3(threads / 256), dim3(256), 0, 0, buf.as<uint32_t>(), 1u, pattern, sink.as<uint32_t>());
  OURO_TRY(hipEventRecord(a, 0));
  hipLaunchKernelGGL(k_atomics, dim3(threads / 256), dim3(256), 0, 0, buf.as<uint32_t>(), iters, pattern, sink.as<uint32_t>());
  OURO_TRY(hipEventRecord(b, 0));
  OURO_TRY(hipEventSynchronize(b));
  float ms = 0;
  OURO_TRY(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ops_per_s = (double)threads * iters / (ms * 1e-3);
  return OURO_OK;
}

int ouro_bench_floor(ouro_handle h, const uint32_t* sizes, uint32_t n, uint32_t block, uint32_t grid, uint32_t tile,
                     uint32_t rounds, double* kernel_ms, uint32_t* grid_out) {
  if (!h || !sizes || !kernel_ms || n == 0 || rounds == 0 || block % 64 != 0 || block == 0 || block > 1024)
    return OURO_E_INVALID;
  if (tile != 1 && tile != 2 && tile != 4) return OURO_E_INVALID;
  IManager& m = *h->impl;
  int cus = 0;
  OURO_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
  const int occ = m.persist_occupancy(block, tile);  // the allocator kernels' limit: the same launch shape
  if (occ <= 0) return OURO_E_INVALID;
  const uint32_t cap = (uint32_t)occ * (uint32_t)cus;
  const uint32_t g = std::min<uint32_t>((n + block * tile - 1) / (block * tile), grid ? std::min(grid, cap) : cap);
  DevBuf d_sizes, d_out;
  OURO_TRY(d_sizes.alloc((size_t)n * 4));
  OURO_TRY(d_out.alloc((size_t)n * 8));
  OURO_TRY(hipMemcpy(d_sizes.p, sizes, (size_t)n * 4, hipMemcpyHostToDevice));
  hipEvent_t e0, e1;
  OURO_TRY(hipEventCreate(&e0));
  OURO_TRY(hipEventCreate(&e1));
  double total = 0;
  hipError_t e = hipSuccess;
  for (uint32_t r = 0; r <= rounds && e == hipSuccess; ++r) {  // launch 0 warms up
    e = m.null_persist(d_sizes.as<uint32_t>(), d_out.as<void*>(), n, 0, block, g, tile, e0, e1);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (r) total += ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return OURO_E_HIP;
  *kernel_ms = total / rounds;
  if (grid_out) *grid_out = g;
  return OURO_OK;
}

int ouro_bench_copy(uint64_t bytes, uint32_t reps, double* gb_per_s) {
  if (!gb_per_s || bytes < 16 || reps == 0) return OURO_E_INVALID;
  const uint64_t n = bytes / 16;
  DevBuf src, dst;
  OURO_TRY(src.alloc(n * 16));
  OURO_TRY(dst.alloc(n * 16));
  OURO_TRY(hipMemset(src.p, 1, n * 16));
  hipEvent_t a, b;
  OURO_TRY(hipEventCreate(&a));
  OURO_TRY(hipEventCreate(&b));
  const uint32_t grid = 256 * 32;
  hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, src.as<uint4>(), dst.as<uint4>(), n);
  OURO_TRY(hipEventRecord(a, 0));
  for (uint32_t r = 0; r < reps; ++r)
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, src.as<uint4>(), dst.as<uint4>(), n);
  OURO_TRY(hipEventRecord(b, 0));
  OURO_TRY(hipEventSynchronize(b));
  float ms = 0;
  OURO_TRY(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *gb_per_s = 2.0 * (double)(n * 16) * reps / (ms * 1e-3) / 1e9;
  return OURO_OK;
}


int ouro_create(int variant, uint64_t pool_bytes, ouro_handle* out) {
  if (!out) return OURO_E_INVALID;
  *out = nullptr;
  IManager* m = make_variant(variant);
  if (!m) return OURO_E_INVALID;
  std::unique_ptr<ouro_manager> h(new (std::nothrow) ouro_manager);
  if (!h) {
    delete m;
    return OURO_E_NOMEM;
  }
  h->impl.reset(m);
  const hipError_t e = m->init(pool_bytes);
  if (e == hipErrorOutOfMemory) return OURO_E_NOMEM;
  if (e == hipErrorInvalidValue) return OURO_E_INVALID;
  if (e != hipSuccess) return OURO_E_HIP;
  *out = h.release();
  return OURO_OK;
}

int ouro_reinitialize(ouro_handle h, double overallocation_factor) {
  if (!h || !(overallocation_factor > 0.0)) return OURO_E_INVALID;
  const hipError_t e = h->impl->reinitialize(overallocation_factor);
  if (e == hipErrorOutOfMemory) return OURO_E_NOMEM;
  if (e == hipErrorInvalidValue) return OURO_E_INVALID;
  return e == hipSuccess ? OURO_OK : OURO_E_HIP;
}

int ouro_destroy(ouro_handle h) {
  delete h;
  return OURO_OK;
}

void* ouro_device_manager(ouro_handle h) { return h ? h->impl->device_manager() : nullptr; }
void* ouro_data_base(ouro_handle h) { return h ? h->impl->data_base() : nullptr; }

int ouro_layout(int variant, uint64_t pool_bytes, uint64_t* max_chunks, uint64_t* allocation_size) {
  std::unique_ptr<IManager> m(make_variant(variant));
  if (!m) return OURO_E_INVALID;
  const ouro::PoolLayout L = m->pool_layout(pool_bytes);
  if (max_chunks) *max_chunks = L.max_chunks;
  if (allocation_size) *allocation_size = L.allocation_size;
  return OURO_OK;
}

int ouro_malloc_batch(ouro_handle h, const uint32_t* d_sizes, void** d_out, uint32_t n, int agg, void* stream) {
  if (!h || (n && (!d_sizes || !d_out))) return OURO_E_INVALID;
  if (agg != 1 && agg != 16 && agg != 64) return OURO_E_INVALID;
  return h->impl->malloc_batch(d_sizes, d_out, n, agg, (hipStream_t)stream, nullptr, 256) == hipSuccess ? OURO_OK
                                                                                                     : OURO_E_HIP;
}

int ouro_free_batch(ouro_handle h, void* const* d_ptrs, uint32_t n, int agg, void* stream) {
  if (!h || (n && !d_ptrs)) return OURO_E_INVALID;
  if (agg != 1 && agg != 16 && agg != 64) return OURO_E_INVALID;
  return h->impl->free_batch(d_ptrs, n, agg, (hipStream_t)stream, 256) == hipSuccess ? OURO_OK : OURO_E_HIP;
}

int ouro_query(ouro_handle h, ouro_state* out) {
  if (!h || !out) return OURO_E_INVALID;
  return h->impl->query(out) == hipSuccess ? OURO_OK : OURO_E_HIP;
}

int ouro_preload(ouro_handle h, uint32_t class_index, uint32_t available_pages, uint32_t* chunk_out) {
  if (!h || !chunk_out || class_index >= h->impl->num_classes()) return OURO_E_INVALID;
  ouro_variant_info g;
  h->impl->geometry(&g);
  // pages per chunk of the class (a class of the second instance spans two of its classes' range)
  const uint32_t page = g.smallest_page << class_index;
  const uint32_t chunk = page <= g.chunk_bytes ? g.chunk_bytes : g.largest_page;
  if (available_pages > chunk / page) return OURO_E_INVALID;
  uint32_t c = 0;
  const hipError_t e = h->impl->preload(class_index, (int)available_pages, &c);
  if (e != hipSuccess) return OURO_E_HIP;
  *chunk_out = c;
  return c == 0xFFFFFFFFu ? OURO_E_NOMEM : OURO_OK;
}

int ouro_set_statistics(ouro_handle h, int on) {
  if (!h) return OURO_E_INVALID;
#if OURO_STATISTICS
  if (hipDeviceSynchronize() != hipSuccess) return OURO_E_HIP;
  const uint32_t flag = on ? 1u : 0u;
  const unsigned long long zero[4] = {0, 0, 0, 0};
  if (on && hipMemcpy(h->impl->device_stats(), zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess)
    return OURO_E_HIP;
  return hipMemcpy(h->impl->device_stats_on(), &flag, 4, hipMemcpyHostToDevice) == hipSuccess ? OURO_OK : OURO_E_HIP;
#else
  (void)on;
  return OURO_E_INVALID;  // built without statistics
#endif
}

int ouro_set_recycling(ouro_handle h, uint32_t min_queued) {
  if (!h) return OURO_E_INVALID;
  ouro_state st;
  if (h->impl->query(&st) != hipSuccess) return OURO_E_HIP;
  if (st.kind[0] != OURO_CQ && st.kind[0] != OURO_VACQ && st.kind[0] != OURO_VLCQ) return OURO_E_INVALID;
  if (hipDeviceSynchronize() != hipSuccess) return OURO_E_HIP;
  return hipMemcpy(h->impl->device_recycle_min(), &min_queued, 4, hipMemcpyHostToDevice) == hipSuccess ? OURO_OK
                                                                                                    : OURO_E_HIP;
}

int ouro_statistics(ouro_handle h, uint64_t* out4) {
  if (!h || !out4) return OURO_E_INVALID;
#if OURO_STATISTICS
  if (hipDeviceSynchronize() != hipSuccess) return OURO_E_HIP;
  return hipMemcpy(out4, h->impl->device_stats(), 32, hipMemcpyDeviceToHost) == hipSuccess ? OURO_OK : OURO_E_HIP;
#else
  return OURO_E_INVALID;
#endif
}

int ouro_snapshot_size(ouro_handle h, uint64_t* bytes) {
  if (!h || !bytes) return OURO_E_INVALID;
  *bytes = h->impl->snapshot_bytes();
  return OURO_OK;
}

int ouro_snapshot(ouro_handle h, void* dst) {
  if (!h || !dst) return OURO_E_INVALID;
  if (hipDeviceSynchronize() != hipSuccess) return OURO_E_HIP;
  return hipMemcpy(dst, h->impl->device_manager(), h->impl->snapshot_bytes(), hipMemcpyDefault) == hipSuccess
             ? OURO_OK
             : OURO_E_HIP;
}

int ouro_restore(ouro_handle h, const void* src) {
  if (!h || !src) return OURO_E_INVALID;
  if (hipDeviceSynchronize() != hipSuccess) return OURO_E_HIP;
  const hipError_t e = h->impl->restore(src);
  return e == hipSuccess ? OURO_OK : (e == hipErrorInvalidValue ? OURO_E_INVALID : OURO_E_HIP);
}

int ouro_set_origin(ouro_handle h, uint32_t origin) {
  if (!h || origin % ouro::kQCSpots) return OURO_E_INVALID;
  const hipError_t e = h->impl->set_origin(origin);
  return e == hipSuccess ? OURO_OK : (e == hipErrorInvalidValue ? OURO_E_INVALID : OURO_E_HIP);
}

int ouro_variant_geometry(int variant, ouro_variant_info* out) {
  if (!out) return OURO_E_INVALID;
  std::unique_ptr<IManager> m(make_variant(variant));
  if (!m) return OURO_E_INVALID;
  m->geometry(out);
  return OURO_OK;
}

int ouro_run_batch(ouro_handle h, const ouro_batch_config* cfg, ouro_batch_result* out) {
  if (!h || !cfg || !out || !cfg->sizes) return OURO_E_INVALID;
  if (cfg->agg != 0 && cfg->agg != 1 && cfg->agg != 16 && cfg->agg != 64) return OURO_E_INVALID;
  const uint32_t block = cfg->block ? cfg->block : 256;
  if (block % 64 != 0 || block > 1024) return OURO_E_INVALID;
  const uint32_t tile = cfg->tile ? cfg->tile : 1;
  if (tile != 1 && tile != 2 && tile != 4) return OURO_E_INVALID;
  if (tile != 1 && cfg->agg != 0) return OURO_E_INVALID;  // tiles are a form of the block-wide kernels
  memset(out, 0, sizeof(*out));
  const uint32_t n = cfg->requests;
  if (n == 0 || (cfg->rounds == 0 && !cfg->check)) return OURO_OK;
  IManager& m = *h->impl;

  DevBuf d_sizes, d_ptrs, d_errs;
  OURO_TRY(d_sizes.alloc((size_t)n * 4));
  OURO_TRY(d_ptrs.alloc((size_t)n * 8));
  OURO_TRY(d_errs.alloc(32));
  OURO_TRY(hipMemcpy(d_sizes.p, cfg->sizes, (size_t)n * 4, hipMemcpyHostToDevice));
  OURO_TRY(hipMemset(d_ptrs.p, 0, (size_t)n * 8));
  OURO_TRY(hipMemset(d_errs.p, 0, 32));

  int cus = 0;
  OURO_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
  uint32_t grid = (n + block - 1) / block;
  if (cfg->agg == 0) {
    grid = (n + block * tile - 1) / (block * tile);
    const int occ = m.persist_occupancy(block, tile);
    if (occ <= 0) return OURO_E_INVALID;  // the block size does not fit the persistent kernels
    // more workgroups than can be resident would serialise behind each other anyway
    const uint32_t cap = (uint32_t)occ * (uint32_t)cus;
    grid = std::min<uint32_t>(grid, cfg->grid ? std::min<uint32_t>(cfg->grid, cap) : cap);
  }
  out->grid = grid;
  const uint8_t* lo = reinterpret_cast<const uint8_t*>(m.data_base());
  const uint8_t* hi = lo + m.max_chunks() * 8320ull;
  ouro_variant_info geo;
  m.geometry(&geo);
  const uint32_t heap_above = geo.largest_page;  // larger requests are served by the device heap
  const uint32_t tgrid = (n + 255) / 256;

  hipStream_t s;
  OURO_TRY(hipStreamCreate(&s));
  hipEvent_t ev[4];
  for (auto& e : ev) OURO_TRY(hipEventCreate(&e));
  void** ptrs = d_ptrs.as<void*>();
  const uint32_t* sizes = d_sizes.as<uint32_t>();
  unsigned long long* errs = d_errs.as<unsigned long long>();
  auto alloc = [&]() -> hipError_t {
    return cfg->agg == 0 ? m.malloc_persist(sizes, ptrs, n, s, block, grid, tile)
                         : m.malloc_batch(sizes, ptrs, n, cfg->agg, s, nullptr, block);
  };
  auto release = [&]() -> hipError_t {
    return cfg->agg == 0 ? m.free_persist(ptrs, n, s, block, grid, tile) : m.free_batch(ptrs, n, cfg->agg, s, block);
  };
  auto touch = [&]() -> hipError_t {
    hipLaunchKernelGGL(k_touch, dim3(tgrid), dim3(256), 0, s, ptrs, sizes, n, lo, hi, 0, errs, heap_above);
    hipLaunchKernelGGL(k_touch, dim3(tgrid), dim3(256), 0, s, ptrs, sizes, n, lo, hi, 1, errs, heap_above);
    return hipGetLastError();
  };
  auto round = [&]() -> hipError_t {
    hipError_t e;
    if ((e = hipEventRecord(ev[0], s)) != hipSuccess || (e = alloc()) != hipSuccess) return e;
    if ((e = hipEventRecord(ev[1], s)) != hipSuccess || (e = touch()) != hipSuccess) return e;
    if ((e = hipEventRecord(ev[2], s)) != hipSuccess || (e = release()) != hipSuccess) return e;
    return hipEventRecord(ev[3], s);
  };

  // validation round: the live set after the alloc kernel must be overlap-free and page-aligned
  if (cfg->check) {
    OURO_TRY(alloc());
    OURO_TRY(hipStreamSynchronize(s));
    std::vector<void*> hp(n);
    OURO_TRY(hipMemcpy(hp.data(), d_ptrs.p, (size_t)n * 8, hipMemcpyDeviceToHost));
    std::vector<uint64_t> offs(n);
    std::vector<uint32_t> sz(cfg->sizes, cfg->sizes + n);
    for (uint32_t i = 0; i < n; ++i)
      offs[i] = hp[i] ? (uint64_t)(reinterpret_cast<const uint8_t*>(hp[i]) - lo) : UINT64_MAX;
    if (cfg->offsets_out) memcpy(cfg->offsets_out, offs.data(), (size_t)n * 8);
    ouro_trace_result tr;
    memset(&tr, 0, sizeof(tr));
    ouro_variant_info gi;
    m.geometry(&gi);
    check_live(offs, sz, &tr, gi);
    out->overlap_errors = tr.overlap_errors;
    out->align_errors = tr.align_errors;
    OURO_TRY(touch());
    OURO_TRY(release());
    OURO_TRY(hipStreamSynchronize(s));
  }
  for (uint32_t w = 0; w < cfg->warmup; ++w) OURO_TRY(round());
  OURO_TRY(hipStreamSynchronize(s));
  // counters cover the timed rounds only (reset outside any captured graph)
  OURO_TRY(hipMemsetAsync(d_errs.p, 0, 32, s));
  OURO_TRY(hipStreamSynchronize(s));

  // graph form: the alloc, touch and free phases are captured as three graphs and replayed one after
  // the other, with the timing events recorded between the launches on the stream (event records
  // captured inside a graph do not time the nodes between them on this runtime)
  hipGraph_t g[3] = {nullptr, nullptr, nullptr};
  hipGraphExec_t ge[3] = {nullptr, nullptr, nullptr};
  auto drop = [&]() {
    for (auto& x : ge)
      if (x) (void)hipGraphExecDestroy(x);
    for (auto& x : g)
      if (x) (void)hipGraphDestroy(x);
  };
  if (cfg->graph && cfg->rounds) {
    for (int k = 0; k < 3; ++k) {
      OURO_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      const hipError_t ce = k == 0 ? alloc() : k == 1 ? touch() : release();
      const hipError_t ee = hipStreamEndCapture(s, &g[k]);
      if (ce != hipSuccess || ee != hipSuccess || hipGraphInstantiate(&ge[k], g[k], nullptr, nullptr, 0) != hipSuccess) {
        drop();
        return OURO_E_HIP;
      }
    }
    out->graph_used = 1;
  }
  auto graph_round = [&]() -> hipError_t {
    hipError_t e;
    for (int k = 0; k < 3; ++k)
      if ((e = hipEventRecord(ev[k], s)) != hipSuccess || (e = hipGraphLaunch(ge[k], s)) != hipSuccess) return e;
    return hipEventRecord(ev[3], s);
  };
  for (uint32_t r = 0; r < cfg->rounds; ++r) {
    const hipError_t re = ge[0] ? graph_round() : round();
    if (re != hipSuccess || hipEventSynchronize(ev[3]) != hipSuccess) {
      drop();
      return OURO_E_HIP;
    }
    float t = 0;
    OURO_TRY(hipEventElapsedTime(&t, ev[0], ev[1]));
    out->alloc_ms += t;
    OURO_TRY(hipEventElapsedTime(&t, ev[1], ev[2]));
    out->touch_ms += t;
    OURO_TRY(hipEventElapsedTime(&t, ev[2], ev[3]));
    out->free_ms += t;
    OURO_TRY(hipEventElapsedTime(&t, ev[0], ev[3]));
    out->round_ms += t;
  }
  OURO_TRY(hipStreamSynchronize(s));
  drop();
  unsigned long long e3[3];
  OURO_TRY(hipMemcpy(e3, d_errs.p, 24, hipMemcpyDeviceToHost));
  if (cfg->kernel_timing && cfg->agg == 0 && cfg->rounds) {
    // kernel-only durations: the same rounds again as plain launches, each persistent kernel timed by
    // the start / end events its own dispatch records
    hipEvent_t ek[4];
    for (auto& e : ek) OURO_TRY(hipEventCreate(&e));
    const uint32_t kr = cfg->rounds;
    for (uint32_t r = 0; r < kr; ++r) {
      OURO_TRY(m.malloc_persist(sizes, ptrs, n, s, block, grid, tile, nullptr, ek[0], ek[1]));
      OURO_TRY(touch());
      OURO_TRY(m.free_persist(ptrs, n, s, block, grid, tile, ek[2], ek[3]));
      OURO_TRY(hipEventSynchronize(ek[3]));
      float t = 0;
      OURO_TRY(hipEventElapsedTime(&t, ek[0], ek[1]));
      out->alloc_kernel_ms += t;
      OURO_TRY(hipEventElapsedTime(&t, ek[2], ek[3]));
      out->free_kernel_ms += t;
    }
    out->kernel_rounds = kr;
    for (auto& e : ek) (void)hipEventDestroy(e);
  }
  if (cfg->latency) {
    // one more round, untimed, sampling the allocation call on lane 0 of every wave
    const uint32_t ns = (n + 63) / 64;
    DevBuf d_lat;
    OURO_TRY(d_lat.alloc((size_t)ns * 8));
    OURO_TRY(hipMemset(d_lat.p, 0, (size_t)ns * 8));
    OURO_TRY(cfg->agg == 0 ? m.malloc_persist(sizes, ptrs, n, s, block, grid, tile, d_lat.as<uint64_t>())
                           : m.malloc_batch(sizes, ptrs, n, cfg->agg, s, d_lat.as<uint64_t>(), block));
    OURO_TRY(touch());
    OURO_TRY(release());
    OURO_TRY(hipStreamSynchronize(s));
    std::vector<uint64_t> lat(ns);
    OURO_TRY(hipMemcpy(lat.data(), d_lat.p, (size_t)ns * 8, hipMemcpyDeviceToHost));
    std::sort(lat.begin(), lat.end());
    int khz = 0;
    (void)hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0);
    const double ns_per_tick = khz > 0 ? 1e6 / (double)khz : 10.0;
    auto pct = [&](double q) { return (double)lat[(size_t)std::min<double>(ns - 1, std::floor(q * (ns - 1) + 0.5))] * ns_per_tick; };
    out->lat_p50_ns = pct(0.50);
    out->lat_p99_ns = pct(0.99);
    out->lat_max_ns = (double)lat.back() * ns_per_tick;
    out->lat_samples = ns;
  }
  out->verify_errors = e3[0];
  out->bounds_errors = e3[1];
  out->null_returns = e3[2];
  out->mallocs = (uint64_t)n * cfg->rounds;
  for (uint32_t i = 0; i < n; ++i) out->touch_bytes += 2ull * (cfg->sizes[i] & ~3u);
  ouro_state st;
  if (m.query(&st) != hipSuccess) return OURO_E_HIP;
  out->error_word = st.error;
  out->peak_chunks = st.next_free_chunk;
  for (auto& e : ev) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(s);
  return OURO_OK;
}

int ouro_run_batch_static(ouro_handle h, const ouro_batch_config* cfg, void* stream, ouro_batch_result* out) {
  if (!h || !cfg || !out || !cfg->sizes) return OURO_E_INVALID;
  if (cfg->agg != 0 && cfg->agg != 1 && cfg->agg != 16 && cfg->agg != 64) return OURO_E_INVALID;
  // graph replay, latency samples and kernel-only timing need events or buffers of their own
  if (cfg->graph || cfg->latency || cfg->kernel_timing) return OURO_E_INVALID;
  if (cfg->requests > kStaticRequests) return OURO_E_INVALID;  // beyond the static scratch
  const uint32_t block = cfg->block ? cfg->block : 256;
  if (block % 64 != 0 || block > 1024) return OURO_E_INVALID;
  const uint32_t tile = cfg->tile ? cfg->tile : 1;
  if (tile != 1 && tile != 2 && tile != 4) return OURO_E_INVALID;
  if (tile != 1 && cfg->agg != 0) return OURO_E_INVALID;
  memset(out, 0, sizeof(*out));
  const uint32_t n = cfg->requests;
  if (n == 0 || (cfg->rounds == 0 && !cfg->check)) return OURO_OK;
  IManager& m = *h->impl;
  uint8_t* sc = nullptr;
  OURO_TRY(static_scratch(&sc));
  unsigned long long* errs = reinterpret_cast<unsigned long long*>(sc);
  uint32_t* sizes = reinterpret_cast<uint32_t*>(sc + kStaticCounterBytes);
  void** ptrs = reinterpret_cast<void**>(sc + kStaticCounterBytes + (size_t)kStaticRequests * 4);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);

  int cus = 0;
  OURO_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
  uint32_t grid = (n + block - 1) / block;
  if (cfg->agg == 0) {
    grid = (n + block * tile - 1) / (block * tile);
    const int occ = m.persist_occupancy(block, tile);
    if (occ <= 0) return OURO_E_INVALID;
    const uint32_t cap = (uint32_t)occ * (uint32_t)cus;
    grid = std::min<uint32_t>(grid, cfg->grid ? std::min<uint32_t>(cfg->grid, cap) : cap);
  }
  out->grid = grid;
  const uint8_t* lo = reinterpret_cast<const uint8_t*>(m.data_base());
  const uint8_t* hi = lo + m.max_chunks() * 8320ull;
  ouro_variant_info geo;
  m.geometry(&geo);
  const uint32_t heap_above = geo.largest_page;  // larger requests are served by the device heap
  const uint32_t tgrid = (n + 255) / 256;
  auto round = [&]() -> hipError_t {
    hipError_t e = cfg->agg == 0 ? m.malloc_persist(sizes, ptrs, n, s, block, grid, tile)
                                 : m.malloc_batch(sizes, ptrs, n, cfg->agg, s, nullptr, block);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_touch, dim3(tgrid), dim3(256), 0, s, ptrs, sizes, n, lo, hi, 0, errs, heap_above);
    hipLaunchKernelGGL(k_touch, dim3(tgrid), dim3(256), 0, s, ptrs, sizes, n, lo, hi, 1, errs, heap_above);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return cfg->agg == 0 ? m.free_persist(ptrs, n, s, block, grid, tile) : m.free_batch(ptrs, n, cfg->agg, s, block);
  };
  OURO_TRY(hipMemcpyAsync(sizes, cfg->sizes, (size_t)n * 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s, errs, 4u);
  OURO_TRY(hipGetLastError());
  if (cfg->check) {
    // the observed round: its blocks (the pointer array keeps them after the free) as offsets
    OURO_TRY(round());
    if (cfg->offsets_out) {
      OURO_TRY(hipMemcpyAsync(cfg->offsets_out, ptrs, (size_t)n * 8, hipMemcpyDeviceToHost, s));
      OURO_TRY(hipStreamSynchronize(s));
      for (uint32_t i = 0; i < n; ++i) {
        const uint64_t p = cfg->offsets_out[i];
        cfg->offsets_out[i] = p ? p - reinterpret_cast<uint64_t>(lo) : UINT64_MAX;
      }
    }
  }
  for (uint32_t w = 0; w < cfg->warmup; ++w) OURO_TRY(round());
  // counters cover the rounds after the observed and warm-up ones, as in ouro_run_batch
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s, errs, 4u);
  OURO_TRY(hipGetLastError());
  for (uint32_t r = 0; r < cfg->rounds; ++r) OURO_TRY(round());
  unsigned long long e3[3] = {0, 0, 0};
  uint32_t words[2] = {0, 0};
  OURO_TRY(hipMemcpyAsync(e3, errs, 24, hipMemcpyDeviceToHost, s));
  OURO_TRY(hipMemcpyAsync(&words[0], m.device_error_word(), 4, hipMemcpyDeviceToHost, s));
  OURO_TRY(hipMemcpyAsync(&words[1], m.device_next_free_chunk(), 4, hipMemcpyDeviceToHost, s));
  OURO_TRY(hipStreamSynchronize(s));
  out->verify_errors = e3[0];
  out->bounds_errors = e3[1];
  out->null_returns = e3[2];
  out->mallocs = (uint64_t)n * cfg->rounds;
  for (uint32_t i = 0; i < n; ++i) out->touch_bytes += 2ull * (cfg->sizes[i] & ~3u);
  out->error_word = words[0];
  out->peak_chunks = words[1];
  return OURO_OK;
}

int ouro_gen_sizes(uint64_t seed, uint32_t n, int dist, double param, uint32_t classes, uint32_t* out) {
  if (!out || classes == 0 || classes > 16) return OURO_E_INVALID;
  if (dist == 0) {
    for (uint32_t i = 0; i < n; ++i) out[i] = (uint32_t)param;
    return OURO_OK;
  }
  if (dist == 1) {
    for (uint32_t i = 0; i < n; ++i) out[i] = 16u << (uint32_t)(keyed(seed, i) % classes);
    return OURO_OK;
  }
  if (dist == 2) {
    // Zipf(param) over class ranks: P(q) proportional to (q + 1)^-param
    std::vector<double> cdf(classes);
    double acc = 0;
    for (uint32_t q = 0; q < classes; ++q) cdf[q] = (acc += std::pow((double)(q + 1), -param));
    for (uint32_t q = 0; q < classes; ++q) cdf[q] /= acc;
    for (uint32_t i = 0; i < n; ++i) {
      const double u = unit(keyed(seed, i));
      uint32_t q = 0;
      while (q + 1 < classes && u >= cdf[q]) ++q;
      out[i] = 16u << q;
    }
    return OURO_OK;
  }
  return OURO_E_INVALID;
}

int ouro_gen_batch(const ouro_batch_workload* w, uint32_t* sizes, uint64_t cap, uint64_t* n_out,
                   uint32_t* requests_out, double* zipf_out) {
  if (!w || !n_out || w->workloads == 0 || w->classes == 0 || w->classes > 16 || w->max_allocs == 0 ||
      w->lo == 0 || w->lo > w->hi || w->s_lo > w->s_hi)
    return OURO_E_INVALID;
  const uint64_t seed = w->seed;
  uint64_t total = 0;
  std::vector<double> cdf(w->classes);
  for (uint32_t k = 0; k < w->workloads; ++k) {
    // request count: lognormal(mu, sigma) by Box-Muller from two keyed uniforms, rounded, clamped
    const double u1 = 1.0 - unit(keyed(seed ^ 0x5851F42D4C957F2Dull, k));
    const double u2 = unit(keyed(seed ^ 0x14057B7EF767814Full, k));
    const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    const double r = std::exp(w->mu + w->sigma * z);
    const uint32_t req = (uint32_t)std::min<double>(w->hi, std::max<double>(w->lo, std::floor(r + 0.5)));
    const double s = w->s_lo + (w->s_hi - w->s_lo) * unit(keyed(seed ^ 0x2545F4914F6CDD1Dull, k));
    if (requests_out) requests_out[k] = req;
    if (zipf_out) zipf_out[k] = s;
    double acc = 0;
    for (uint32_t q = 0; q < w->classes; ++q) cdf[q] = (acc += std::pow((double)(q + 1), -s));
    for (uint32_t q = 0; q < w->classes; ++q) cdf[q] /= acc;
    const uint64_t ws = seed ^ (0x9E3779B97F4A7C15ull * (k + 1));
    for (uint32_t i = 0; i < req; ++i) {
      const uint32_t na = 1u + (uint32_t)(keyed(ws, i) % w->max_allocs);
      for (uint32_t j = 0; j < na; ++j, ++total) {
        if (!sizes || total >= cap) continue;
        const double u = unit(keyed(ws ^ 0xD6E8FEB86659FD93ull, (uint64_t)i * 4u + j));
        uint32_t q = 0;
        while (q + 1 < w->classes && u >= cdf[q]) ++q;
        sizes[total] = 16u << q;
      }
    }
  }
  *n_out = total;
  return OURO_OK;
}

int ouro_run_trace(ouro_handle h, const ouro_trace_config* cfg, ouro_trace_result* out) {
  if (!h || !cfg || !out) return OURO_E_INVALID;
  if (cfg->agg != 1 && cfg->agg != 16 && cfg->agg != 64) return OURO_E_INVALID;
  if (!cfg->sizes && !cfg->round_sizes) return OURO_E_INVALID;
  IManager& m = *h->impl;
  const uint32_t n = cfg->threads, R = cfg->rounds;
  const uint32_t block = cfg->block ? cfg->block : 256;
  if (block % 64 != 0 || block > 1024) return OURO_E_INVALID;
  memset(out, 0, sizeof(*out));
  if (n == 0 || R == 0) return OURO_OK;

  DevBuf d_sizes[2], d_ptrs[2], d_skip_cur, d_skip_prev, d_errs, d_lat;
  for (int i = 0; i < 2; ++i) {
    OURO_TRY(d_sizes[i].alloc((size_t)n * 4));
    OURO_TRY(d_ptrs[i].alloc((size_t)n * 8));
    OURO_TRY(hipMemset(d_ptrs[i].p, 0, (size_t)n * 8));
  }
  OURO_TRY(d_skip_cur.alloc(n));
  OURO_TRY(d_skip_prev.alloc(n));
  OURO_TRY(d_errs.alloc(32));
  OURO_TRY(hipMemset(d_errs.p, 0, 32));
  const uint32_t nlat = (n + 63) / 64;
  if (cfg->latency) OURO_TRY(d_lat.alloc((size_t)nlat * 8));

  hipStream_t s;
  OURO_TRY(hipStreamCreate(&s));
  hipEvent_t ev[6];
  for (auto& e : ev) OURO_TRY(hipEventCreate(&e));

  const uint8_t* lo = reinterpret_cast<const uint8_t*>(m.data_base());
  const uint8_t* hi = lo + m.max_chunks() * 8320ull;
  ouro_variant_info geo;
  m.geometry(&geo);
  const uint32_t heap_above = geo.largest_page;  // larger requests are served by the device heap
  const uint32_t grid = (n + block - 1) / block;
  std::vector<uint32_t> hsizes(n), psizes(n);
  std::vector<uint8_t> hold(n, 0), phold(n, 0), inv(n, 0);
  std::vector<uint64_t> offs(n), poffs(n, UINT64_MAX);
  std::vector<void*> hptrs(n);
  std::vector<uint64_t> lat_all;
  int wallclock_khz = 0;
  (void)hipDeviceGetAttribute(&wallclock_khz, hipDeviceAttributeWallClockRate, 0);
  const bool held = cfg->hold_permille != 0;

  for (uint32_t r = 0; r < R; ++r) {
    const int cur = r & 1, prv = cur ^ 1;
    const uint32_t* src = cfg->round_sizes ? cfg->round_sizes + (size_t)r * n : cfg->sizes;
    memcpy(hsizes.data(), src, (size_t)n * 4);
    OURO_TRY(hipMemcpyAsync(d_sizes[cur].p, hsizes.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
    if (held) {
      // a held block of round r is freed in round r + 1 (or after the last round)
      for (uint32_t i = 0; i < n; ++i)
        hold[i] = (keyed(cfg->hold_seed ^ (0x9E3779B9ull * (r + 1)), i) % 1000) < cfg->hold_permille;
      for (uint32_t i = 0; i < n; ++i) inv[i] = !phold[i];
      OURO_TRY(hipMemcpyAsync(d_skip_cur.p, hold.data(), n, hipMemcpyHostToDevice, s));
      OURO_TRY(hipMemcpyAsync(d_skip_prev.p, inv.data(), n, hipMemcpyHostToDevice, s));
      OURO_TRY(hipStreamSynchronize(s));
    }
    OURO_TRY(hipEventRecord(ev[0], s));
    OURO_TRY(m.malloc_batch(d_sizes[cur].as<uint32_t>(), d_ptrs[cur].as<void*>(), n, cfg->agg, s,
                            cfg->latency ? d_lat.as<uint64_t>() : nullptr, block));
    OURO_TRY(hipEventRecord(ev[1], s));
    hipLaunchKernelGGL(k_touch, dim3(grid), dim3(block), 0, s, d_ptrs[cur].as<void*>(), d_sizes[cur].as<uint32_t>(), n,
                       lo, hi, 0, d_errs.as<unsigned long long>(), heap_above);
    OURO_TRY(hipEventRecord(ev[2], s));
    hipLaunchKernelGGL(k_touch, dim3(grid), dim3(block), 0, s, d_ptrs[cur].as<void*>(), d_sizes[cur].as<uint32_t>(), n,
                       lo, hi, 1, d_errs.as<unsigned long long>(), heap_above);
    OURO_TRY(hipEventRecord(ev[3], s));
    OURO_TRY(hipStreamSynchronize(s));
    {
      OURO_TRY(hipMemcpy(hptrs.data(), d_ptrs[cur].p, (size_t)n * 8, hipMemcpyDeviceToHost));
      for (uint32_t i = 0; i < n; ++i) {
        offs[i] = hptrs[i] ? (uint64_t)(reinterpret_cast<const uint8_t*>(hptrs[i]) - lo) : UINT64_MAX;
        out->null_returns += hptrs[i] == nullptr;
      }
      if (cfg->offsets_out) memcpy(cfg->offsets_out + (size_t)r * n, offs.data(), (size_t)n * 8);
    }
    if (cfg->check_rounds) {
      // live set = this round's blocks + the blocks held back from the previous round
      std::vector<uint64_t> live = offs;
      std::vector<uint32_t> lsz = hsizes;
      if (held)
        for (uint32_t i = 0; i < n; ++i)
          if (phold[i] && poffs[i] != UINT64_MAX) {
            live.push_back(poffs[i]);
            lsz.push_back(psizes[i]);
          }
      ouro_variant_info gi;
      m.geometry(&gi);
      check_live(live, lsz, out, gi);
    }
    OURO_TRY(hipEventRecord(ev[4], s));
    if (held) {
      if (r > 0) OURO_TRY(m.free_masked(d_ptrs[prv].as<void*>(), d_skip_prev.as<uint8_t>(), n, cfg->agg, s, block));
      OURO_TRY(m.free_masked(d_ptrs[cur].as<void*>(), d_skip_cur.as<uint8_t>(), n, cfg->agg, s, block));
    } else {
      OURO_TRY(m.free_batch(d_ptrs[cur].as<void*>(), n, cfg->agg, s, block));
    }
    OURO_TRY(hipEventRecord(ev[5], s));
    OURO_TRY(hipEventSynchronize(ev[5]));
    float t = 0;
    OURO_TRY(hipEventElapsedTime(&t, ev[0], ev[1]));
    out->alloc_ms += t;
    OURO_TRY(hipEventElapsedTime(&t, ev[1], ev[2]));
    out->write_ms += t;
    OURO_TRY(hipEventElapsedTime(&t, ev[2], ev[3]));
    out->verify_ms += t;
    OURO_TRY(hipEventElapsedTime(&t, ev[4], ev[5]));
    out->free_ms += t;
    out->mallocs += n;
    if (cfg->latency) {
      std::vector<uint64_t> l(nlat);
      OURO_TRY(hipMemcpy(l.data(), d_lat.p, (size_t)nlat * 8, hipMemcpyDeviceToHost));
      lat_all.insert(lat_all.end(), l.begin(), l.end());
    }
    for (uint32_t i = 0; i < n; ++i) out->touch_bytes += hptrs[i] ? 2ull * (hsizes[i] & ~3u) : 0ull;
    phold.swap(hold);
    psizes.swap(hsizes);
    poffs.swap(offs);
  }
  if (held) {
    // the last round's held blocks
    const int last = (R - 1) & 1;
    for (uint32_t i = 0; i < n; ++i) inv[i] = !phold[i];
    OURO_TRY(hipMemcpy(d_skip_prev.p, inv.data(), n, hipMemcpyHostToDevice));
    OURO_TRY(hipEventRecord(ev[4], s));
    OURO_TRY(m.free_masked(d_ptrs[last].as<void*>(), d_skip_prev.as<uint8_t>(), n, cfg->agg, s, block));
    OURO_TRY(hipEventRecord(ev[5], s));
    OURO_TRY(hipEventSynchronize(ev[5]));
    float t = 0;
    OURO_TRY(hipEventElapsedTime(&t, ev[4], ev[5]));
    out->free_ms += t;
  }
  OURO_TRY(hipStreamSynchronize(s));
  unsigned long long errs[2];
  OURO_TRY(hipMemcpy(errs, d_errs.p, 16, hipMemcpyDeviceToHost));
  out->verify_errors = errs[0];
  out->bounds_errors = errs[1];
  ouro_state st;
  if (m.query(&st) != hipSuccess) return OURO_E_HIP;
  out->error_word = st.error;
  out->peak_chunks = st.next_free_chunk;
  out->frees = out->mallocs - out->null_returns;
  if (!lat_all.empty()) {
    std::sort(lat_all.begin(), lat_all.end());
    const double ns_per_tick = wallclock_khz > 0 ? 1e6 / (double)wallclock_khz : 10.0;
    auto pct = [&](double p) {
      size_t i = (size_t)std::min<double>((double)lat_all.size() - 1, std::floor(p * (double)(lat_all.size() - 1) + 0.5));
      return (double)lat_all[i] * ns_per_tick;
    };
    out->lat_samples = lat_all.size();
    out->lat_p50_ns = pct(0.50);
    out->lat_p99_ns = pct(0.99);
    out->lat_max_ns = (double)lat_all.back() * ns_per_tick;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(s);
  return OURO_OK;
}

// ------------------------------------------------------------------ quiescent invariant check
int ouro_check_quiescent(ouro_handle h, ouro_check_result* out) {
  if (!h || !out) return OURO_E_INVALID;
  memset(out, 0, sizeof(*out));
  out->first_class = -1;
  out->first_chunk = -1;
  IManager& m = *h->impl;
  ouro_state st;
  if (m.query(&st) != hipSuccess) return OURO_E_HIP;
  auto fail = [&](int cls, int64_t chunk, const char* msg, long long a, long long b) {
    if (out->violations == 0) {
      out->first_class = cls;
      out->first_chunk = (int32_t)chunk;
      snprintf(out->message, sizeof(out->message), "%s (%lld vs %lld)", msg, a, b);
    }
    out->violations++;
  };
  const uint8_t* d_data = reinterpret_cast<const uint8_t*>(m.data_base());
  const uint8_t* d_mem = reinterpret_cast<const uint8_t*>(m.device_manager());
  const uint32_t nfc = std::min<uint32_t>(st.next_free_chunk, (uint32_t)st.max_chunks);
  // chunk headers and the run-start locator
  std::vector<uint8_t> hdrs((size_t)nfc * 128 + 128);
  if (nfc) OURO_TRY(hipMemcpy2D(hdrs.data(), 128, d_data, 8320, 128, nfc, hipMemcpyDeviceToHost));
  const uint64_t loc_bytes = ((st.max_chunks + 31) / 32) * 4;
  std::vector<uint32_t> loc((loc_bytes + 3) / 4);
  {
    const uint64_t loc_region = ((loc_bytes + 8319) / 8320) * 8320;
    OURO_TRY(hipMemcpy(loc.data(), d_mem + st.allocation_size - loc_region, loc_bytes, hipMemcpyDeviceToHost));
  }
  const bool chunk_kind = st.kind[0] == OURO_CQ || st.kind[0] == OURO_VACQ || st.kind[0] == OURO_VLCQ;
  std::vector<int64_t> class_pages(st.num_classes, 0), class_chunks(st.num_classes, 0);
  std::vector<std::vector<uint32_t>> members(st.num_classes);
  for (uint32_t c = 0; c < nfc; ++c) {
    if (!((loc[c >> 5] >> (c & 31)) & 1u)) continue;
    const uint32_t* w = reinterpret_cast<const uint32_t*>(hdrs.data() + (size_t)c * 128);
    const uint32_t ps = w[0];
    if (ps == 0) continue;  // queue chunk or recycled queue chunk
    int k = -1;
    for (uint32_t q = 0; q < st.num_classes; ++q)
      if (st.page_size[q] == ps) k = (int)q;
    if (k < 0) {
      fail(-1, c, "chunk header holds an unknown page size", ps, 0);
      continue;
    }
    out->chunks_checked++;
    const uint32_t ppc = st.pages_per_chunk[k];
    class_chunks[k]++;
    members[k].push_back(c);
    if (chunk_kind) {
      const ouro::ChunkHdr* ch = reinterpret_cast<const ouro::ChunkHdr*>(w);
      if (ch->size != ppc) fail(k, c, "chunk size field differs from pages per chunk", ch->size, ppc);
      if (ch->count != (int32_t)ppc) fail(k, c, "chunk free count is not full", ch->count, ppc);
      for (uint32_t wi = 0; wi < ouro::kMaskWords; ++wi) {
        const uint32_t lo = wi * 64;
        uint64_t full = 0;
        if (ppc > lo) full = (ppc - lo >= 64) ? ~0ull : ((1ull << (ppc - lo)) - 1ull);
        if (ch->mask[wi] != full) fail(k, c, "chunk availability mask is not full", (long long)wi, 0);
      }
      class_pages[k] += ch->count;
      // every page is free again, so the chunk must sit in its class queue inside [front_, back_)
      if (!ouro::qp_queued(ch->queue_pos) || ch->queue_pos != ouro::qp_at(ch->gen, (uint32_t)ch->queue_pos) ||
          (uint32_t)ch->queue_pos - st.front[k] >= st.back[k] - st.front[k])
        fail(k, c, "chunk with free pages is not inside its queue window", (long long)ch->queue_pos, st.front[k]);
    } else {
      const ouro::PageChunkHdr* ph = reinterpret_cast<const ouro::PageChunkHdr*>(w);
      if (ph->number_pages != ppc) fail(k, c, "page chunk page count differs", ph->number_pages, ppc);
      class_pages[k] += ppc;
    }
  }
  for (uint32_t k = 0; k < st.num_classes; ++k) {
    out->pages_free += class_pages[k];
    if (st.sem_count[k] != class_pages[k]) fail(k, -1, "semaphore count != free pages of the class", st.sem_count[k], class_pages[k]);
    if (st.sem_expected[k] != 0) fail(k, -1, "semaphore expected != 0", st.sem_expected[k], 0);
    if (st.sem_reserved[k] != 0) fail(k, -1, "semaphore reserved != 0", st.sem_reserved[k], 0);
    const uint32_t kind = st.kind[k];
    const bool ck = kind == OURO_CQ || kind == OURO_VACQ || kind == OURO_VLCQ;
    if (ck) {
      if ((int64_t)st.carved[k] != class_chunks[k]) fail(k, -1, "carved chunks != chunks of the class", st.carved[k], class_chunks[k]);
      if ((int64_t)st.queue_count[k] != class_chunks[k]) fail(k, -1, "queue count != chunks of the class", st.queue_count[k], class_chunks[k]);
    } else {
      if ((int64_t)st.carved[k] != class_pages[k]) fail(k, -1, "carved pages != pages of the class", st.carved[k], class_pages[k]);
      if ((int64_t)(uint32_t)(st.back[k] - st.front[k]) != class_pages[k])
        fail(k, -1, "queue back - front != free pages", (long long)(uint32_t)(st.back[k] - st.front[k]), class_pages[k]);
    }
    // standard array queues: the live ring window must hold exactly the class's pages / chunks
    uint32_t* ring = nullptr;
    uint32_t rsize = 0;
    m.class_ring(k, &ring, &rsize);
    if ((kind == OURO_PQ || kind == OURO_CQ) && ring && rsize) {
      const uint32_t len = st.back[k] - st.front[k];
      if (len > rsize) {
        fail(k, -1, "queue window exceeds ring", len, rsize);
        continue;
      }
      std::vector<uint32_t> win(len);
      const uint32_t f = st.front[k] & (rsize - 1);
      const uint32_t first = std::min(len, rsize - f);
      if (first) OURO_TRY(hipMemcpy(win.data(), ring + f, (size_t)first * 4, hipMemcpyDeviceToHost));
      if (len > first) OURO_TRY(hipMemcpy(win.data() + first, ring, (size_t)(len - first) * 4, hipMemcpyDeviceToHost));
      std::unordered_set<uint64_t> seen;
      std::unordered_set<uint32_t> mem(members[k].begin(), members[k].end());
      for (uint32_t i = 0; i < len; ++i) {
        const uint32_t v = win[i];
        if (v == ouro::kMarker) {
          fail(k, -1, "unwritten slot inside the queue window", i, len);
          continue;
        }
        if (kind == OURO_CQ && v == ouro::kTomb) continue;  // removed chunk
        const uint32_t c = kind == OURO_PQ ? (v >> 9) : v;
        const uint32_t pg = kind == OURO_PQ ? (v & 511u) : 0u;
        if (!mem.count(c)) fail(k, c, "queue entry names a chunk outside the class", c, 0);
        if (kind == OURO_CQ && c < nfc) {
          const ouro::ChunkHdr* ch = reinterpret_cast<const ouro::ChunkHdr*>(hdrs.data() + (size_t)c * 128);
          if (ch->queue_pos != ouro::qp_at(ch->gen, st.front[k] + i))
            fail(k, c, "chunk queue_pos differs from its queue position", (long long)ch->queue_pos, st.front[k] + i);
        }
        if (kind == OURO_PQ && pg >= st.pages_per_chunk[k]) fail(k, c, "queue entry page out of range", pg, st.pages_per_chunk[k]);
        if (!seen.insert(v).second) fail(k, c, "duplicate queue entry", v, 0);
      }
      if (kind == OURO_CQ && (int64_t)seen.size() != class_chunks[k])
        fail(k, -1, "chunk queue entries != chunks of the class", (long long)seen.size(), class_chunks[k]);
    }
  }
  if (st.error) fail(-1, -1, "device error word is set", st.error, 0);
  return OURO_OK;
}

}  // extern "C"
