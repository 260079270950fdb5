// Type-erased wrapper around one Ouroboros<...> instantiation, used by the C ABI.
// Each variant is compiled in its own translation unit (variant.hip) so the six instantiations
// build in parallel.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <memory>
#include <vector>

#include "ouro/instances.hpp"
#include "ouro_capi.h"

namespace ouro_lib {

struct ZipfTable;
struct RequestBlock;

// Free capacity of a pool: `shared` base chunks any class can take (the bump region left and the
// reuse ring) and, per class in chain order, its free pages (the class semaphore's count).
struct PoolFree {
  uint64_t shared;
  int64_t pages[16];
  uint32_t page_size[16];
  uint32_t ppc[16];
  uint32_t add[16];  // base chunks one chunk of the class spans
  double cost[16];   // base chunks a fresh chunk of the class takes, with its share of queue chunks
  uint32_t classes;
};

struct IManager {
  virtual ~IManager() {}
  virtual hipError_t init(uint64_t pool_bytes) = 0;
  virtual void* device_manager() = 0;
  virtual void* data_base() = 0;
  virtual uint64_t allocation_size() = 0;
  virtual uint64_t snapshot_bytes() = 0;
  virtual hipError_t restore(const void* src) = 0;
  virtual hipError_t reinitialize(double factor) = 0;
  virtual uint64_t max_chunks() = 0;
  virtual int variant() = 0;
  virtual hipError_t malloc_batch(const uint32_t* d_sizes, void** d_out, uint32_t n, int agg, hipStream_t s,
                                  uint64_t* d_lat, uint32_t block) = 0;
  virtual hipError_t free_batch(void* const* d_ptrs, uint32_t n, int agg, hipStream_t s, uint32_t block) = 0;
  virtual hipError_t free_masked(void* const* d_ptrs, const uint8_t* d_skip, uint32_t n, int agg, hipStream_t s,
                                 uint32_t block) = 0;
  // Batched form: persistent kernels over a flat request list with block-wide aggregation
  // (Ouroboros::malloc_tile / free_tile, `tile` = 1, 2 or 4 requests per thread). grid = persistent
  // workgroups.
  // e0 / e1 (optional): events the kernel's own dispatch records at its start and end
  // (hipExtLaunchKernel), for kernel-only durations outside any captured graph.
  virtual hipError_t malloc_persist(const uint32_t* d_sizes, void** d_out, uint32_t n, hipStream_t s, uint32_t block,
                                    uint32_t grid, uint32_t tile, uint64_t* d_lat = nullptr,
                                    hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) = 0;
  virtual hipError_t free_persist(void* const* d_ptrs, uint32_t n, hipStream_t s, uint32_t block, uint32_t grid,
                                  uint32_t tile, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) = 0;
  // the floor kernel (k_null_persist) at the same launch shape, timed by its own dispatch's events
  virtual hipError_t null_persist(const uint32_t* d_sizes, void** d_out, uint32_t n, hipStream_t s, uint32_t block,
                                  uint32_t grid, uint32_t tile, hipEvent_t e0, hipEvent_t e1) = 0;
  // workgroups of `block` threads one CU holds at once for the persistent kernels of `tile` (the
  // smallest over the three)
  virtual int persist_occupancy(uint32_t block, uint32_t tile) = 0;
  // Serving form: one launch of `blocks` 256-thread workgroups described by d_desc (k_request /
  // k_request_tile, `tile` request threads per thread for agg 0) on stream s.
  virtual hipError_t launch_requests(const struct RequestBlock* d_desc, uint32_t blocks, const struct ZipfTable& z,
                                     unsigned long long* out, int agg, hipStream_t s, uint32_t tile = 1) = 0;
  // What the pool can still hand out, read from a host copy of the manager state (state_bytes()
  // bytes at device_manager()).
  virtual void pool_free(const void* state, struct PoolFree* out) = 0;
  virtual size_t state_bytes() = 0;
  virtual hipError_t query(ouro_state* out) = 0;
  virtual void geometry(ouro_variant_info* out) = 0;
  virtual ouro::PoolLayout pool_layout(uint64_t pool_bytes) = 0;
  // immutable queue placement of class k (chain order)
  virtual void class_ring(uint32_t k, uint32_t** ring, uint32_t* ring_size) = 0;
  // fresh pool only: every class queue starts at position `origin` (a multiple of 2048)
  virtual hipError_t set_origin(uint32_t origin) = 0;
  // Ouroboros::preloadChunk(k, available) on one device thread; *chunk = the chunk, or kMarker
  virtual hipError_t preload(uint32_t k, int available, uint32_t* chunk) = 0;
  virtual uint32_t num_classes() = 0;
  // device addresses of the sticky error word and next_free_chunk inside the pool's manager state
  virtual uint32_t* device_error_word() = 0;
  virtual uint32_t* device_stats_on() = 0;
  virtual uint32_t* device_recycle_min() = 0;
  virtual unsigned long long* device_stats() = 0;
  virtual uint32_t* device_next_free_chunk() = 0;
#ifdef OURO_DIAG
  virtual hipError_t diag_read(unsigned long long* out16, int reset) = 0;
#endif
};

// Host trace randomness: SplitMix64, an independent stream per (seed, index).
inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
inline uint64_t keyed(uint64_t seed, uint64_t i) { return splitmix64(splitmix64(seed) ^ (i * 0xD1B54A32D192ED03ull)); }
inline double unit(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

IManager* make_variant(int v);

// Factories, one per variant translation unit.
IManager* make_pq();
IManager* make_cq();
IManager* make_vapq();
IManager* make_vacq();
IManager* make_vlpq();
IManager* make_vlcq();
IManager* make_geo6();
IManager* make_geo7();
IManager* make_geo8();
IManager* make_geo9();
IManager* make_geo10();
IManager* make_geo11();

// ------------------------------------------------------------------ harness kernels
template <class M, int W>
__global__ void __launch_bounds__(256) k_malloc(M* mm, const uint32_t* __restrict__ sizes, void** __restrict__ out,
                                                uint32_t n) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= n) return;
  out[tid] = mm->template malloc<W>(sizes[tid]);
}

// Latency sample: lane 0 of every wave records wall_clock64() ticks around its malloc.
template <class M, int W>
__global__ void __launch_bounds__(256) k_malloc_lat(M* mm, const uint32_t* __restrict__ sizes, void** __restrict__ out,
                                                    uint32_t n, uint64_t* __restrict__ lat) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= n) return;
  const uint64_t t0 = wall_clock64();
  void* p = mm->template malloc<W>(sizes[tid]);
  const uint64_t t1 = wall_clock64();
  out[tid] = p;
  if ((tid & 63u) == 0) lat[tid >> 6] = t1 - t0;
}

template <class M, int W>
__global__ void __launch_bounds__(256) k_free(M* mm, void* const* __restrict__ ptrs, uint32_t n) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= n) return;
  mm->template free<W>(ptrs[tid]);
}

template <class M, int W>
__global__ void __launch_bounds__(256) k_free_masked(M* mm, void* const* __restrict__ ptrs,
                                                     const uint8_t* __restrict__ skip, uint32_t n) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= n) return;
  mm->template free<W>(skip[tid] ? nullptr : ptrs[tid]);
}

// Persistent block-aggregated kernels: each workgroup walks tiles of T * blockDim.x requests (request
// t * blockDim.x + i of a tile goes to thread i; the loop bound is uniform per block, as malloc_tile /
// free_tile synchronise the block). The next tile's input is loaded while the current tile is served.
template <class M, uint32_t T>
__global__ void __launch_bounds__(1024) k_malloc_persist(M* mm, const uint32_t* __restrict__ sizes,
                                                         void** __restrict__ out, uint32_t n) {
  __shared__ typename M::template TileScratch<T> sm;
  const uint32_t tile = T * blockDim.x, step = gridDim.x * tile;
  uint32_t base = blockIdx.x * tile;
  size_t size[T];
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    const uint32_t i = base + t * blockDim.x + threadIdx.x;
    size[t] = i < n ? sizes[i] : 0u;
  }
  for (; base < n; base += step) {
    uint32_t next[T];
#pragma unroll
    for (uint32_t t = 0; t < T; ++t) {
      const uint32_t j = base + step + t * blockDim.x + threadIdx.x;
      next[t] = j < n ? sizes[j] : 0u;
    }
    void* p[T];
    mm->template malloc_tile<T>(size, p, sm);
#pragma unroll
    for (uint32_t t = 0; t < T; ++t) {
      const uint32_t i = base + t * blockDim.x + threadIdx.x;
      if (i < n) out[i] = p[t];
      size[t] = next[t];
    }
  }
}

// k_malloc_persist's launch, loop and loads around null_tile: the floor an allocator kernel of this
// shape cannot go below.
template <class M, uint32_t T>
__global__ void __launch_bounds__(1024) k_null_persist(M* mm, const uint32_t* __restrict__ sizes,
                                                       void** __restrict__ out, uint32_t n) {
  __shared__ typename M::template TileScratch<T> sm;
  const uint32_t tile = T * blockDim.x, step = gridDim.x * tile;
  uint32_t base = blockIdx.x * tile;
  size_t size[T];
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    const uint32_t i = base + t * blockDim.x + threadIdx.x;
    size[t] = i < n ? sizes[i] : 0u;
  }
  for (; base < n; base += step) {
    uint32_t next[T];
#pragma unroll
    for (uint32_t t = 0; t < T; ++t) {
      const uint32_t j = base + step + t * blockDim.x + threadIdx.x;
      next[t] = j < n ? sizes[j] : 0u;
    }
    void* p[T];
    mm->template null_tile<T>(size, p, sm);
#pragma unroll
    for (uint32_t t = 0; t < T; ++t) {
      const uint32_t i = base + t * blockDim.x + threadIdx.x;
      if (i < n) out[i] = p[t];
      size[t] = next[t];
    }
  }
}

// k_malloc_persist with the wall-clock time of each malloc_tile call recorded on lane 0 of every
// wave that holds a request (sample of request i at lat[i / 64]; the T requests of a wave's lanes
// share one call).
template <class M, uint32_t T>
__global__ void __launch_bounds__(1024) k_malloc_persist_lat(M* mm, const uint32_t* __restrict__ sizes,
                                                             void** __restrict__ out, uint32_t n,
                                                             uint64_t* __restrict__ lat) {
  __shared__ typename M::template TileScratch<T> sm;
  const uint32_t tile = T * blockDim.x, step = gridDim.x * tile;
  for (uint32_t base = blockIdx.x * tile; base < n; base += step) {
    size_t size[T];
#pragma unroll
    for (uint32_t t = 0; t < T; ++t) {
      const uint32_t i = base + t * blockDim.x + threadIdx.x;
      size[t] = i < n ? sizes[i] : 0u;
    }
    void* p[T];
    const uint64_t t0 = wall_clock64();
    mm->template malloc_tile<T>(size, p, sm);
    const uint64_t t1 = wall_clock64();
#pragma unroll
    for (uint32_t t = 0; t < T; ++t) {
      const uint32_t i = base + t * blockDim.x + threadIdx.x;
      if (i < n) {
        out[i] = p[t];
        if ((threadIdx.x & 63u) == 0) lat[i >> 6] = t1 - t0;
      }
    }
  }
}

template <class M, uint32_t T>
__global__ void __launch_bounds__(1024) k_free_persist(M* mm, void* const* __restrict__ ptrs, uint32_t n) {
  __shared__ typename M::template TileScratch<T> sm;
  const uint32_t tile = T * blockDim.x, step = gridDim.x * tile;
  uint32_t base = blockIdx.x * tile;
  void* ptr[T];
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    const uint32_t i = base + t * blockDim.x + threadIdx.x;
    ptr[t] = i < n ? ptrs[i] : nullptr;
  }
  for (; base < n; base += step) {
    void* next[T];
#pragma unroll
    for (uint32_t t = 0; t < T; ++t) {
      const uint32_t j = base + step + t * blockDim.x + threadIdx.x;
      next[t] = j < n ? ptrs[j] : nullptr;
    }
    mm->template free_tile<T>(ptr, sm);
#pragma unroll
    for (uint32_t t = 0; t < T; ++t) ptr[t] = next[t];
  }
}

inline uint32_t grid_for(uint32_t n, uint32_t block) { return (n + block - 1) / block; }

// One preload (Ouroboros::preloadChunk) by a single device thread.
template <class M>
__global__ void k_preload(M* mm, uint32_t k, int available, uint32_t* out) {
  if (threadIdx.x == 0) out[0] = mm->preloadChunk(k, available);
}

// ------------------------------------------------------------------ serving requests
// Zipf draw over the class ranks: class q is the first with r < thr[q] (r a 32-bit random number).
struct ZipfTable {
  uint32_t thr[16];
  uint32_t classes;
};

__device__ __forceinline__ uint64_t dev_splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ uint64_t dev_keyed(uint64_t seed, uint64_t i) {
  return dev_splitmix64(dev_splitmix64(seed) ^ (i * 0xD1B54A32D192ED03ull));
}

// Work of one workgroup of a serving launch: threads first .. first + blockDim.x of request `req`
// (n threads, pattern seed `seed`). One launch carries the blocks of every request it batches.
struct RequestBlock {
  uint64_t seed;
  uint32_t req;
  uint32_t first;
  uint32_t n;
  uint32_t pad;
};

// Thread i of a request: a block of class q (16 B << q, q drawn from the top bits of keyed(seed, i)
// against the Zipf table), filled with a pattern keyed by (seed, i), read back and freed.
__device__ __forceinline__ void request_draw(const RequestBlock& d, uint32_t i, const ZipfTable& z, uint32_t& q,
                                             uint32_t& v0) {
  const uint64_t key = dev_keyed(d.seed, i);
  const uint32_t r = (uint32_t)(key >> 32);
  q = 0;
  while (q + 1 < z.classes && r >= z.thr[q]) ++q;
  v0 = (uint32_t)key;
}

__device__ __forceinline__ void request_touch(uint32_t* p, uint32_t q, uint32_t v0, unsigned long long& ok,
                                              unsigned long long& bad) {
  const uint32_t words = (16u << q) / 4u;
  for (uint32_t k = 0; k < words; k += 4)  // blocks are at least 16-byte aligned
    *reinterpret_cast<uint4*>(p + k) = make_uint4(v0 + k, v0 + k + 1, v0 + k + 2, v0 + k + 3);
  for (uint32_t k = 0; k < words; k += 4) {
    const uint4 v = *reinterpret_cast<const uint4*>(p + k);
    const uint32_t g = (v.x == v0 + k) + (v.y == v0 + k + 1) + (v.z == v0 + k + 2) + (v.w == v0 + k + 3);
    ok += g;
    bad += 4u - g;
  }
}

// Per-request result line: out[4 * req + 0] += words read back intact, + 1 words that were not,
// + 2 null returns. A request's result depends only on its seed and size, never on what ran beside it.
__device__ __forceinline__ void request_report(unsigned long long* out, uint32_t req, unsigned long long ok,
                                               unsigned long long bad, unsigned long long nul) {
  for (int off = 32; off > 0; off >>= 1) {
    ok += __shfl_xor(ok, off);
    bad += __shfl_xor(bad, off);
    nul += __shfl_xor(nul, off);
  }
  if (__lane_id() == 0) {
    if (ok) atomicAdd(&out[4 * (size_t)req + 0], ok);
    if (bad) atomicAdd(&out[4 * (size_t)req + 1], bad);
    if (nul) atomicAdd(&out[4 * (size_t)req + 2], nul);
  }
}

// A serving launch through the per-thread device malloc of width W.
template <class M, int W>
__global__ void __launch_bounds__(256) k_request(M* mm, const RequestBlock* __restrict__ desc, ZipfTable z,
                                                 unsigned long long* __restrict__ out) {
  const RequestBlock d = desc[blockIdx.x];
  const uint32_t i = d.first + threadIdx.x;
  unsigned long long ok = 0, bad = 0, nul = 0;
  if (i < d.n) {
    uint32_t q, v0;
    request_draw(d, i, z, q, v0);
    uint32_t* p = reinterpret_cast<uint32_t*>(mm->template malloc<W>(16u << q));
    if (p) {
      request_touch(p, q, v0, ok, bad);
      mm->template free<W>(p);
    } else {
      nul = 1;
    }
  }
  request_report(out, d.req, ok, bad, nul);
}

// The same launch served block-wide (Ouroboros::malloc_tile / free_tile): every thread of the block
// calls together and serves T request threads, request thread first + t * blockDim.x + i on thread
// i; a thread past the request's end passes size 0 / nullptr.
template <class M, uint32_t T>
__global__ void __launch_bounds__(1024) k_request_tile(M* mm, const RequestBlock* __restrict__ desc, ZipfTable z,
                                                       unsigned long long* __restrict__ out) {
  __shared__ typename M::template TileScratch<T> sm;
  const RequestBlock d = desc[blockIdx.x];
  unsigned long long ok = 0, bad = 0, nul = 0;
  size_t size[T];
  uint32_t q[T], v0[T];
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    const uint32_t i = d.first + t * blockDim.x + threadIdx.x;
    q[t] = 0;
    v0[t] = 0;
    size[t] = 0;
    if (i < d.n) {
      request_draw(d, i, z, q[t], v0[t]);
      size[t] = 16u << q[t];
    }
  }
  void* p[T];
  mm->template malloc_tile<T>(size, p, sm);
#pragma unroll
  for (uint32_t t = 0; t < T; ++t) {
    if (p[t]) request_touch(reinterpret_cast<uint32_t*>(p[t]), q[t], v0[t], ok, bad);
    else if (size[t]) ++nul;
  }
  mm->template free_tile<T>(p, sm);
  request_report(out, d.req, ok, bad, nul);
}

#ifdef OURO_DIAG
// diagnostic builds: dynamic LDS per block (bytes) from OURO_DIAG_LDS, to cap the occupancy
inline size_t launch_lds() {
  const char* e = getenv("OURO_DIAG_LDS");
  return e ? (size_t)atoi(e) : 0;
}
#else
inline size_t launch_lds() { return 0; }
#endif

template <class M>
struct Manager final : IManager {
  M mm;
  int var;
  explicit Manager(int v) : var(v) {}
  hipError_t init(uint64_t pool_bytes) override { return mm.initialize(pool_bytes); }
  void* device_manager() override { return mm.getDeviceMemoryManager(); }
  void* data_base() override { return mm.memory.d_data; }
  uint64_t allocation_size() override { return mm.memory.allocation_size; }
  uint64_t snapshot_bytes() override { return mm.snapshot_bytes(); }
  hipError_t restore(const void* src) override { return mm.restore(src); }
  hipError_t reinitialize(double factor) override { return mm.reinitialize((float)factor); }
  uint64_t max_chunks() override { return mm.memory.max_chunks; }
  int variant() override { return var; }

  template <int W>
  hipError_t launch_malloc(const uint32_t* d_sizes, void** d_out, uint32_t n, hipStream_t s, uint64_t* d_lat,
                           uint32_t block) {
    M* dm = mm.getDeviceMemoryManager();
    if (d_lat)
      hipLaunchKernelGGL((k_malloc_lat<M, W>), dim3(grid_for(n, block)), dim3(block), launch_lds(), s, dm, d_sizes, d_out, n,
                         d_lat);
    else
      hipLaunchKernelGGL((k_malloc<M, W>), dim3(grid_for(n, block)), dim3(block), launch_lds(), s, dm, d_sizes, d_out, n);
    return hipGetLastError();
  }
  hipError_t malloc_batch(const uint32_t* d_sizes, void** d_out, uint32_t n, int agg, hipStream_t s, uint64_t* d_lat,
                          uint32_t block) override {
    if (n == 0) return hipSuccess;
    switch (agg) {
      case 1: return launch_malloc<1>(d_sizes, d_out, n, s, d_lat, block);
      case 16: return launch_malloc<16>(d_sizes, d_out, n, s, d_lat, block);
      case 64: return launch_malloc<64>(d_sizes, d_out, n, s, d_lat, block);
      default: return hipErrorInvalidValue;
    }
  }
  template <int W>
  hipError_t launch_free(void* const* d_ptrs, const uint8_t* d_skip, uint32_t n, hipStream_t s, uint32_t block) {
    M* dm = mm.getDeviceMemoryManager();
    if (d_skip)
      hipLaunchKernelGGL((k_free_masked<M, W>), dim3(grid_for(n, block)), dim3(block), launch_lds(), s, dm, d_ptrs, d_skip, n);
    else
      hipLaunchKernelGGL((k_free<M, W>), dim3(grid_for(n, block)), dim3(block), launch_lds(), s, dm, d_ptrs, n);
    return hipGetLastError();
  }
  template <uint32_t T>
  hipError_t launch_malloc_persist(const uint32_t* d_sizes, void** d_out, uint32_t n, hipStream_t s, uint32_t block,
                                   uint32_t grid, uint64_t* d_lat, hipEvent_t e0, hipEvent_t e1) {
    M* dm = mm.getDeviceMemoryManager();
    if (d_lat)
      hipLaunchKernelGGL((k_malloc_persist_lat<M, T>), dim3(grid), dim3(block), 0, s, dm, d_sizes, d_out, n, d_lat);
    else if (e0 || e1)
      hipExtLaunchKernelGGL((k_malloc_persist<M, T>), dim3(grid), dim3(block), 0, s, e0, e1, 0, dm, d_sizes, d_out, n);
    else
      hipLaunchKernelGGL((k_malloc_persist<M, T>), dim3(grid), dim3(block), 0, s, dm, d_sizes, d_out, n);
    return hipGetLastError();
  }
  hipError_t malloc_persist(const uint32_t* d_sizes, void** d_out, uint32_t n, hipStream_t s, uint32_t block,
                            uint32_t grid, uint32_t tile, uint64_t* d_lat, hipEvent_t e0, hipEvent_t e1) override {
    if (n == 0) return hipSuccess;
    switch (tile) {
      case 1: return launch_malloc_persist<1>(d_sizes, d_out, n, s, block, grid, d_lat, e0, e1);
      case 2: return launch_malloc_persist<2>(d_sizes, d_out, n, s, block, grid, d_lat, e0, e1);
      case 4: return launch_malloc_persist<4>(d_sizes, d_out, n, s, block, grid, d_lat, e0, e1);
      default: return hipErrorInvalidValue;
    }
  }
  hipError_t null_persist(const uint32_t* d_sizes, void** d_out, uint32_t n, hipStream_t s, uint32_t block,
                          uint32_t grid, uint32_t tile, hipEvent_t e0, hipEvent_t e1) override {
    M* dm = mm.getDeviceMemoryManager();
    switch (tile) {
      case 1: hipExtLaunchKernelGGL((k_null_persist<M, 1>), dim3(grid), dim3(block), 0, s, e0, e1, 0, dm, d_sizes, d_out, n); break;
      case 2: hipExtLaunchKernelGGL((k_null_persist<M, 2>), dim3(grid), dim3(block), 0, s, e0, e1, 0, dm, d_sizes, d_out, n); break;
      case 4: hipExtLaunchKernelGGL((k_null_persist<M, 4>), dim3(grid), dim3(block), 0, s, e0, e1, 0, dm, d_sizes, d_out, n); break;
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  template <uint32_t T>
  hipError_t launch_free_persist(void* const* d_ptrs, uint32_t n, hipStream_t s, uint32_t block, uint32_t grid,
                                 hipEvent_t e0, hipEvent_t e1) {
    M* dm = mm.getDeviceMemoryManager();
    if (e0 || e1)
      hipExtLaunchKernelGGL((k_free_persist<M, T>), dim3(grid), dim3(block), 0, s, e0, e1, 0, dm, d_ptrs, n);
    else
      hipLaunchKernelGGL((k_free_persist<M, T>), dim3(grid), dim3(block), 0, s, dm, d_ptrs, n);
    return hipGetLastError();
  }
  hipError_t free_persist(void* const* d_ptrs, uint32_t n, hipStream_t s, uint32_t block, uint32_t grid,
                          uint32_t tile, hipEvent_t e0, hipEvent_t e1) override {
    if (n == 0) return hipSuccess;
    switch (tile) {
      case 1: return launch_free_persist<1>(d_ptrs, n, s, block, grid, e0, e1);
      case 2: return launch_free_persist<2>(d_ptrs, n, s, block, grid, e0, e1);
      case 4: return launch_free_persist<4>(d_ptrs, n, s, block, grid, e0, e1);
      default: return hipErrorInvalidValue;
    }
  }
  hipError_t launch_requests(const RequestBlock* d_desc, uint32_t blocks, const ZipfTable& z, unsigned long long* out,
                             int agg, hipStream_t s, uint32_t tile) override {
    if (blocks == 0) return hipSuccess;
    M* dm = mm.getDeviceMemoryManager();
    if (agg == 0 && tile != 1 && tile != 2 && tile != 4) return hipErrorInvalidValue;
    switch (agg) {
      case 0:
        if (tile == 4) hipLaunchKernelGGL((k_request_tile<M, 4>), dim3(blocks), dim3(256), 0, s, dm, d_desc, z, out);
        else if (tile == 2) hipLaunchKernelGGL((k_request_tile<M, 2>), dim3(blocks), dim3(256), 0, s, dm, d_desc, z, out);
        else hipLaunchKernelGGL((k_request_tile<M, 1>), dim3(blocks), dim3(256), 0, s, dm, d_desc, z, out);
        break;
      case 1: hipLaunchKernelGGL((k_request<M, 1>), dim3(blocks), dim3(256), 0, s, dm, d_desc, z, out); break;
      case 16: hipLaunchKernelGGL((k_request<M, 16>), dim3(blocks), dim3(256), 0, s, dm, d_desc, z, out); break;
      case 64: hipLaunchKernelGGL((k_request<M, 64>), dim3(blocks), dim3(256), 0, s, dm, d_desc, z, out); break;
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  template <int N, class... Is>
  static void free_pages_of(const ouro::Chain<>&, PoolFree*) {}
  template <int N, class I, class... Rest>
  static void free_pages_of(const ouro::Chain<I, Rest...>& ch, PoolFree* out) {
    for (uint32_t q = 0; q < I::kNQ && out->classes < 16; ++q, ++out->classes) {
      const uint32_t k = out->classes;
      out->pages[k] = ouro::Sem::count(ch.inst.q[q].sem);
      out->page_size[k] = (1u << I::kSmallLog2) << q;
      out->ppc[k] = (I::kChunkSize >> I::kSmallLog2) >> q;
      out->add[k] = I::kAdd;
      // a virtualized queue stores 2048 entries per queue chunk: one per page (page queues) or one per
      // chunk (chunk queues)
      const double entries = is_std(I::kKind) ? 0.0 : I::kChunks ? 1.0 : (double)out->ppc[k];
      out->cost[k] = I::kAdd + entries / ouro::kQCSpots;
    }
    free_pages_of<N + 1, Rest...>(ch.next, out);
  }
  size_t state_bytes() override { return sizeof(M); }
  void pool_free(const void* state, PoolFree* out) override {
    const M* h = reinterpret_cast<const M*>(state);
    const uint64_t limit = (uint64_t)h->memory.max_chunks - h->memory.qc_reserve;
    const int64_t reuse = h->chain.inst.reuse.count;
    out->shared = (limit > h->memory.next_free_chunk ? limit - h->memory.next_free_chunk : 0) + (reuse > 0 ? reuse : 0);
    out->classes = 0;
    free_pages_of<0>(h->chain, out);
  }
  template <uint32_t T>
  static int occupancy_of(uint32_t block) {
    int a = 0, b = 0, c = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_malloc_persist<M, T>, (int)block, 0) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_free_persist<M, T>, (int)block, 0) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, k_malloc_persist_lat<M, T>, (int)block, 0) != hipSuccess) return 0;
    b = b < c ? b : c;
    return a < b ? a : b;
  }
  int persist_occupancy(uint32_t block, uint32_t tile) override {
    switch (tile) {
      case 1: return occupancy_of<1>(block);
      case 2: return occupancy_of<2>(block);
      case 4: return occupancy_of<4>(block);
      default: return 0;
    }
  }
  hipError_t free_batch(void* const* d_ptrs, uint32_t n, int agg, hipStream_t s, uint32_t block) override {
    return free_masked(d_ptrs, nullptr, n, agg, s, block);
  }
  hipError_t free_masked(void* const* d_ptrs, const uint8_t* d_skip, uint32_t n, int agg, hipStream_t s,
                         uint32_t block) override {
    if (n == 0) return hipSuccess;
    switch (agg) {
      case 1: return launch_free<1>(d_ptrs, d_skip, n, s, block);
      case 16: return launch_free<16>(d_ptrs, d_skip, n, s, block);
      case 64: return launch_free<64>(d_ptrs, d_skip, n, s, block);
      default: return hipErrorInvalidValue;
    }
  }

  template <int N, class... Is>
  static void collect(ouro::Chain<>&, ouro_state*, uint32_t&) {}
  template <int N, class I, class... Rest>
  static void collect(ouro::Chain<I, Rest...>& ch, ouro_state* out, uint32_t& k) {
    if (N < 2) out->reuse_count[N] = ch.inst.reuse.count;
    for (uint32_t i = 0; i < I::kNQ && k < OURO_MAX_CLASSES; ++i, ++k) {
      const ouro::ClassQueue& q = ch.inst.q[i];
      out->page_size[k] = q.page_size;
      out->pages_per_chunk[k] = q.ppc;
      out->sem_count[k] = ouro::Sem::count(q.sem);
      out->sem_expected[k] = (uint32_t)ouro::Sem::expected(q.sem);
      out->sem_reserved[k] = 0;  // no reserved field: claims wait on the queue, not the word
      out->front[k] = q.front;
      out->back[k] = q.back;
      out->queue_count[k] = q.count;
      out->carved[k] = q.carved;
      out->front_qc[k] = q.front_qc;
      out->kind[k] = (uint32_t)I::kKind;
    }
    collect<N + 1, Rest...>(ch.next, out, k);
  }

  template <int N, class... Is>
  static void ring_of(ouro::Chain<>&, uint32_t, uint32_t**, uint32_t*) {}
  template <int N, class I, class... Rest>
  static void ring_of(ouro::Chain<I, Rest...>& ch, uint32_t k, uint32_t** ring, uint32_t* size) {
    if (k < I::kNQ) {
      *ring = ch.inst.q[k].ring;
      *size = ch.inst.q[k].ring_size;
      return;
    }
    ring_of<N + 1, Rest...>(ch.next, k - I::kNQ, ring, size);
  }
  void class_ring(uint32_t k, uint32_t** ring, uint32_t* ring_size) override {
    *ring = nullptr;
    *ring_size = 0;
    ring_of<0>(mm.chain, k, ring, ring_size);
  }

  // Moves the fresh class queues of a host copy of the manager state to `origin`: front, back and
  // the page cursor, and the initial queue chunk of a VA / VL queue (its virtual start; VA: its
  // pointer-ring entry moves to the entry of origin's queue chunk).
  template <int N, class... Is>
  static hipError_t origin_of(ouro::Chain<>&, const ouro::Memory&, uint32_t) { return hipSuccess; }
  template <int N, class I, class... Rest>
  static hipError_t origin_of(ouro::Chain<I, Rest...>& ch, const ouro::Memory& mem, uint32_t origin) {
    using namespace ouro;
    for (uint32_t i = 0; i < I::kNQ; ++i) {
      ClassQueue& q = ch.inst.q[i];
      if (q.front != 0 || q.back != 0 || q.sem != Sem::kInit) return hipErrorInvalidValue;
      q.front = q.back = origin;
      q.scan = (uint64_t)origin * q.ppc;
      uint32_t c = kMarker;
      hipError_t e = hipSuccess;
      if (is_va(I::kKind)) {
        const uint32_t id = (origin / kQCSpots) & (I::kQueueEntries - 1), marker = kMarker;
        e = hipMemcpy(&c, q.ring, 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(q.ring, &marker, 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(q.ring + id, &c, 4, hipMemcpyHostToDevice);
      } else if (is_vl(I::kKind)) {
        // the queue chunk's directory entry moves with its virtual start
        c = q.front_qc;
        uint64_t* dir = reinterpret_cast<uint64_t*>(ch.inst.tab.ring[i]);
        const uint64_t empty = ouro::detail::kVLDirEmpty, entry = ((uint64_t)origin << 32) | c;
        e = hipMemcpy(dir, &empty, 8, hipMemcpyHostToDevice);
        if (e == hipSuccess)
          e = hipMemcpy(dir + ((origin / kQCSpots) & (ouro::detail::kVLDir - 1)), &entry, 8, hipMemcpyHostToDevice);
      }
      if (e == hipSuccess && c != kMarker) {
        if (c >= mem.max_chunks) return hipErrorInvalidValue;
        e = hipMemcpy(mem.d_data + (uint64_t)c * kStride + offsetof(QueueChunkHdr, virtual_start), &origin, 4,
                      hipMemcpyHostToDevice);
      }
      if (e != hipSuccess) return e;
    }
    return origin_of<N + 1, Rest...>(ch.next, mem, origin);
  }
  uint32_t* device_field(const void* host_field) {
    const size_t off = (size_t)(reinterpret_cast<const uint8_t*>(host_field) - reinterpret_cast<const uint8_t*>(&mm));
    return reinterpret_cast<uint32_t*>(mm.memory.d_memory + off);
  }
  uint32_t* device_error_word() override { return device_field(&mm.memory.error); }
  uint32_t* device_stats_on() override { return device_field(&mm.memory.stats_on); }
  uint32_t* device_recycle_min() override { return device_field(&mm.memory.recycle_min); }
  unsigned long long* device_stats() override {
    return reinterpret_cast<unsigned long long*>(device_field(&mm.memory.stats[0]));
  }
  uint32_t num_classes() override { return M::kNumClasses; }
  hipError_t preload(uint32_t k, int available, uint32_t* chunk) override {
    uint32_t* d = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&d), 4);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_preload<M>), dim3(1), dim3(1), 0, 0, mm.getDeviceMemoryManager(), k, available, d);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(chunk, d, 4, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return e;
  }
  uint32_t* device_next_free_chunk() override { return device_field(&mm.memory.next_free_chunk); }
  hipError_t set_origin(uint32_t origin) override {
    if (origin % ouro::kQCSpots) return hipErrorInvalidValue;
    std::vector<uint8_t> buf(sizeof(M));
    hipError_t e = hipMemcpy(buf.data(), mm.memory.d_memory, sizeof(M), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    M* h = reinterpret_cast<M*>(buf.data());
    e = origin_of<0>(h->chain, h->memory, origin);
    if (e == hipSuccess) e = hipMemcpy(mm.memory.d_memory, buf.data(), sizeof(M), hipMemcpyHostToDevice);
    return e;
  }

#ifdef OURO_DIAG
  // each variant translation unit holds its own copy of the diagnostic counters
  hipError_t diag_read(unsigned long long* out16, int reset) override {
    hipError_t e = hipMemcpyFromSymbol(out16, HIP_SYMBOL(ouro::g_diag), sizeof(unsigned long long) * 96);
    if (e == hipSuccess && reset) {
      unsigned long long z[96] = {0};
      e = hipMemcpyToSymbol(HIP_SYMBOL(ouro::g_diag), z, sizeof(z));
    }
    return e;
  }
#endif

  ouro::PoolLayout pool_layout(uint64_t pool_bytes) override { return M::layout(pool_bytes); }
  void geometry(ouro_variant_info* out) override {
    using I = decltype(mm.chain.inst);
    memset(out, 0, sizeof(*out));
    out->kind = (uint32_t)I::kKind;
    out->classes = M::kNumClasses;
    out->smallest_page = I::kSmallest;
    out->largest_page = M::kLargestPage;
    out->chunk_bytes = I::kChunkSize;
    out->queue_entries = I::kQueueEntries;
    out->page_capacity = ouro::is_vl(I::kKind) ? 0xFFFFFFFFull
                         : ouro::is_va(I::kKind) ? (uint64_t)I::kQueueEntries * ouro::kQCSpots - 2ull * ouro::kQCSpots
                                                 : (uint64_t)I::kQueueEntries;
  }
  hipError_t query(ouro_state* out) override {
    // the manager state is ~8 KiB; stage it in heap memory, it must not be constructed
    std::vector<uint8_t> buf(sizeof(M));
    hipError_t e = hipMemcpy(buf.data(), mm.memory.d_memory, sizeof(M), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    M* h = reinterpret_cast<M*>(buf.data());
    memset(out, 0, sizeof(*out));
    out->allocation_size = h->memory.allocation_size;
    out->max_chunks = h->memory.max_chunks;
    out->data_offset = (uint64_t)(h->memory.d_data - h->memory.d_memory);
    out->next_free_chunk = h->memory.next_free_chunk;
    out->error = h->memory.error;
    out->variant = (uint32_t)var;
    uint32_t k = 0;
    collect<0>(h->chain, out, k);
    out->num_classes = k;
    return hipSuccess;
  }
};

}  // namespace ouro_lib

// The C handle behind ouro_handle.
struct ouro_manager {
  std::unique_ptr<ouro_lib::IManager> impl;
};
