// A user's kernels written only against the device API (include/ouro/instances.hpp) and the
// reference's names (src/main.cu:10-18,55-63 calls mm->malloc / mm->free the same way): the drop-in
// contract, with malloc and free called from data-dependent branches of a wave. Test code, built
// into tests/dropin/libdropin.so and driven by tests/test_gpu_dropin.py.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <vector>

#include "ouro/instances.hpp"

namespace {

__device__ uint32_t size_of(uint32_t tid, uint32_t round) { return 16u << ((tid * 7u + round) % 12u); }
__device__ uint32_t mark_of(uint32_t tid, uint32_t round) { return tid ^ (round << 24) ^ 0x5A5A0000u; }

// Round r: a thread whose tid + r is odd allocates in the `if` branch and fills its block; the
// others, in the `else` branch, check and free the block they allocated in round r - 1. Every wave
// runs both branches with complementary lanes.
template <typename MemoryManagerType>
__global__ void d_alternate(MemoryManagerType* mm, uint32_t** slots, uint32_t n, uint32_t round,
                            unsigned long long* errs) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= n) return;
  if (((tid + round) & 1u) == 1u) {
    const uint32_t size = size_of(tid, round);
    uint32_t* p = reinterpret_cast<uint32_t*>(mm->malloc(size));
    slots[tid] = p;
    if (p == nullptr) {
      atomicAdd(&errs[0], 1ull);
    } else {
      for (uint32_t k = 0; k < size / 4; ++k) p[k] = mark_of(tid, round);
    }
  } else {
    uint32_t* p = slots[tid];
    if (p != nullptr && round > 0) {
      const uint32_t size = size_of(tid, round - 1);
      for (uint32_t k = 0; k < size / 4; ++k)
        if (p[k] != mark_of(tid, round - 1)) atomicAdd(&errs[1], 1ull);
      mm->free(p);
      slots[tid] = nullptr;
    }
  }
}

// Whatever is still held after the last round.
template <typename MemoryManagerType>
__global__ void d_drain(MemoryManagerType* mm, uint32_t** slots, uint32_t n) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid < n && slots[tid] != nullptr) {
    mm->free(slots[tid]);
    slots[tid] = nullptr;
  }
}

// Every third thread allocates; of those, odd and even ones free from different branches after a
// per-lane loop of data-dependent length.
template <typename MemoryManagerType>
__global__ void d_every_third(MemoryManagerType* mm, uint32_t n, uint32_t seed, unsigned long long* errs) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= n) return;
  if (tid % 3 == 0) {
    const uint32_t size = 16u << ((tid ^ seed) % 10u);
    uint32_t* p = reinterpret_cast<uint32_t*>(mm->malloc(size));
    if (p == nullptr) {
      atomicAdd(&errs[0], 1ull);
      return;
    }
    for (uint32_t k = 0; k < size / 4; ++k) p[k] = tid + seed;
    unsigned long long bad = 0;
    for (uint32_t k = 0; k < size / 4; ++k) bad += p[k] != tid + seed;
    if ((tid / 3) & 1u) {
      if (bad) atomicAdd(&errs[1], bad);
      mm->free(p);
    } else {
      mm->free(p);
      if (bad) atomicAdd(&errs[1], bad);
    }
  }
}

template <typename M>
hipError_t run(M* dm, uint32_t threads, uint32_t rounds, unsigned long long* errs_host) {
  uint32_t** slots = nullptr;
  unsigned long long* errs = nullptr;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&slots), (size_t)threads * sizeof(void*));
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&errs), 2 * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMemset(slots, 0, (size_t)threads * sizeof(void*));
  if (e == hipSuccess) e = hipMemset(errs, 0, 2 * sizeof(unsigned long long));
  const dim3 grid((threads + 255) / 256), block(256);
  for (uint32_t r = 0; r < rounds && e == hipSuccess; ++r) {
    hipLaunchKernelGGL((d_alternate<M>), grid, block, 0, 0, dm, slots, threads, r, errs);
    hipLaunchKernelGGL((d_every_third<M>), grid, block, 0, 0, dm, threads, r * 977u, errs);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL((d_drain<M>), grid, block, 0, 0, dm, slots, threads);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(errs_host, errs, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost);
  if (slots) (void)hipFree(slots);
  if (errs) (void)hipFree(errs);
  return e;
}

}  // namespace

extern "C" {

// The kernels above on a pool created through the C ABI (ouro_device_manager of a handle of variant
// `kind`, 0..5): errs[0] = null returns, errs[1] = words read back wrong. 0 on success.
int dropin_run(int kind, void* device_manager, uint32_t threads, uint32_t rounds, unsigned long long* errs) {
  hipError_t e;
  switch (kind) {
    case 0: e = run(reinterpret_cast<ouro::MultiOuroPQ*>(device_manager), threads, rounds, errs); break;
    case 1: e = run(reinterpret_cast<ouro::MultiOuroCQ*>(device_manager), threads, rounds, errs); break;
    case 2: e = run(reinterpret_cast<ouro::MultiOuroVAPQ*>(device_manager), threads, rounds, errs); break;
    case 3: e = run(reinterpret_cast<ouro::MultiOuroVACQ*>(device_manager), threads, rounds, errs); break;
    case 4: e = run(reinterpret_cast<ouro::MultiOuroVLPQ*>(device_manager), threads, rounds, errs); break;
    case 5: e = run(reinterpret_cast<ouro::MultiOuroVLCQ*>(device_manager), threads, rounds, errs); break;
    default: return 1;
  }
  return e == hipSuccess ? 0 : 2;
}

// The reference's own host sequence (src/main.cu:134-136): a MultiOuroVACQ object, initialize(size),
// getDeviceMemoryManager() passed to the kernels. *error_word = the pool's sticky error bits after
// the run, *next_free = its bump pointer.
int dropin_own_manager(uint64_t pool_bytes, uint32_t threads, uint32_t rounds, unsigned long long* errs,
                       uint32_t* error_word, uint32_t* next_free) {
  using MemoryManagerType = ouro::MultiOuroVACQ;
  MemoryManagerType memory_manager;
  if (memory_manager.initialize(pool_bytes) != hipSuccess) return 2;
  hipError_t e = run(memory_manager.getDeviceMemoryManager(), threads, rounds, errs);
  if (e != hipSuccess) return 2;
  std::vector<uint8_t> state(sizeof(MemoryManagerType));  // raw bytes: a second object would free the pool
  e = hipMemcpy(state.data(), memory_manager.getDeviceMemoryManager(), state.size(), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return 2;
  const auto* s = reinterpret_cast<const MemoryManagerType*>(state.data());
  *error_word = s->memory.error;
  *next_free = s->memory.next_free_chunk;
  return 0;
}

}  // extern "C"
