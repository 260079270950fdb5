// Device primitives for the gfx950 (CDNA4) Ouroboros allocator.
//
// Replaces the reference's inline-PTX helpers (include/Utility.cuh:6-95):
//   ld.global.cg / st.global.cg  -> relaxed agent-scope atomic loads/stores. On gfx950 these
//                                   carry the sc1 bit, i.e. they are coherent across the eight
//                                   XCD-private L2 caches, which a plain load is not.
//   __nanosleep back-off         -> s_sleep with bounded exponential back-off; every spin has
//                                   an iteration cap that records ERR_TIMEOUT instead of hanging.
//   __match_any_sync / __ffs /   -> 64-lane ballot peeling (match<W>), mbcnt-style ranks and
//   __popc / __shfl_sync            ds_bpermute broadcasts on 64-bit exec masks.
//   __threadfence(_block)        -> order(): wait until this lane's device-coherent stores and
//                                   atomics have completed. All allocator metadata is written
//                                   with sc1 stores or atomics, so completion implies agent-scope
//                                   visibility without an L2 write-back (buffer_wbl2).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace ouro {

// Sticky error bits, OR-ed into Memory::error. The reference has ErrorCodes
// (include/Utility.h:259-267) but never sets them; here every failure path records one.
enum : uint32_t {
  ERR_OOM = 1u << 0,          // no chunk left in the pool (Ouroboros.cuh:133 returns false)
  ERR_QUEUE_FULL = 1u << 1,   // a class queue would exceed its capacity (PageQueue_impl.cuh:34,77)
  ERR_VSTART = 1u << 2,       // queue chunk virtual start mismatch (PageQueueVA_impl.cuh:241-246)
  ERR_TIMEOUT = 1u << 3,      // a bounded spin gave up (QueueChunk_impl.cuh:53-58 traps instead)
  ERR_BAD_FREE = 1u << 4,     // free() of a pointer that is not a live page of this pool
  ERR_TOO_LARGE = 1u << 5,    // request above the largest class of the last instance
  ERR_REUSE_FULL = 1u << 6,   // chunk reuse ring overflow (Queue_impl.cuh:35 returns false)
  ERR_LOCATOR = 1u << 7,      // chunk locator walked below index 0 (ChunkLocator.cuh:42-47)
  ERR_CORRUPT = 1u << 8,      // a queue entry named a chunk or page outside the pool
};

constexpr uint32_t kMarker = 0xFFFFFFFFu;  // empty slot, DeletionMarker<index_t> (Definitions.h:31-35)
constexpr uint32_t kTomb = 0xFFFFFFFDu;    // chunk-queue slot whose chunk was removed (front_ may pass it)
constexpr uint32_t kNullPage = 0xFFFFFFFCu;  // page-queue entry standing for a page the pool could not supply
constexpr uint32_t kWave = 64;             // CDNA wavefront width (WARP_SIZE=32 in Definitions.h:16)
// Spin budget: a wait that lasts longer than this (wall-clock ticks of s_memrealtime, 100 MHz on
// MI355X, so about 2 s) gives up and records ERR_TIMEOUT; nested waits are bounded by depth x budget.
constexpr uint64_t kSpinTicks = 200000000ull;

// Statistics counters (Memory::stats) are compiled in unless OURO_STATISTICS is 0, and counted while
// the pool's stats_on flag is set: one atomic per wave (or workgroup) and call, never per request.
#ifndef OURO_STATISTICS
#define OURO_STATISTICS 1
#endif
enum : uint32_t { STAT_PAGE_ALLOC = 0, STAT_PAGE_FREE = 1, STAT_CHUNK_ALLOC = 2, STAT_CHUNK_REUSE = 3 };

#ifdef OURO_DIAG
// Diagnostic event counters (diagnostic builds only).
// [0, 16) counters, [16] events logged, [17, 96) the first 79 event words.
static __device__ unsigned long long g_diag[96];
#define OURO_DIAG_ADD(i, n) atomicAdd(&::ouro::g_diag[i], (unsigned long long)(n))
#define OURO_DIAG_LOG(v)                                                     \
  do {                                                                       \
    const unsigned long long _i = atomicAdd(&::ouro::g_diag[16], 1ull);      \
    if (_i < 79) atomicExch(&::ouro::g_diag[17 + _i], (unsigned long long)(v)); \
  } while (0)
#else
#define OURO_DIAG_ADD(i, n) ((void)0)
#define OURO_DIAG_LOG(v) ((void)0)
#endif

namespace dev {

// A freed block's contents may still sit dirty in the L2 of the XCD that wrote them; each of the
// eight XCDs has its own L2, and agent-scope atomics do not write those lines back. If the block
// were handed to a wave on another XCD first, a later write-back of the old lines would overwrite
// the new owner's data. free() therefore releases at agent scope (buffer_wbl2 sc1, then waits)
// before it publishes a block, which CUDA's device-coherent L2 makes unnecessary on the reference's
// hardware. A new owner needs no acquire: it writes a block before it reads it back.
__device__ __forceinline__ void release_block_contents() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }

// ------------------------------------------------------------------ agent-scope atomics
template <class T>
__device__ __forceinline__ T ld(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T, class U>
__device__ __forceinline__ void st(T* p, U v) {
  __hip_atomic_store(p, (T)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T, class U>
__device__ __forceinline__ T add(T* p, U v) {
  return __hip_atomic_fetch_add(p, (T)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T, class U>
__device__ __forceinline__ T sub(T* p, U v) {
  return __hip_atomic_fetch_sub(p, (T)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T, class U>
__device__ __forceinline__ T and_(T* p, U v) {
  return __hip_atomic_fetch_and(p, (T)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T, class U>
__device__ __forceinline__ T or_(T* p, U v) {
  return __hip_atomic_fetch_or(p, (T)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T, class U>
__device__ __forceinline__ T max_(T* p, U v) {
  return __hip_atomic_fetch_max(p, (T)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T, class U>
__device__ __forceinline__ T xchg(T* p, U v) {
  return __hip_atomic_exchange(p, (T)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Returns the value found at *p (== expected on success), like CUDA's atomicCAS.
template <class T, class U, class V>
__device__ __forceinline__ T cas(T* p, U expected_in, V desired_in) {
  T expected = (T)expected_in;
  const T desired = (T)desired_in;
  __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return expected;
}

// Wait until every earlier memory operation of this lane has completed. Used between
// publishing metadata (chunk headers, queue-chunk slots) and the store or atomic that makes it
// reachable, so a reader on another XCD never finds the handle before the data behind it.
__device__ __forceinline__ void order() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Bounded exponential back-off (Utility.cuh:64-74 uses __nanosleep(10*factor)).
__device__ __forceinline__ void backoff(uint32_t iter) {
  if (iter < 4) return;
  if (iter < 64) __builtin_amdgcn_s_sleep(1);
  else __builtin_amdgcn_s_sleep(8);
}

// A bounded wait: the first 256 iterations are free, after that the wall clock decides. Once any
// wait of the kernel has given up (ERR_TIMEOUT in the pool's error word), every other wait gives up
// at its next clock check, so one stuck wait cannot keep thousands of waves spinning.
struct Spin {
  uint32_t it = 0;
  uint64_t t0 = 0;
  const uint32_t* err = nullptr;
  __device__ __forceinline__ Spin() {}
  __device__ __forceinline__ explicit Spin(const uint32_t* error_word) : err(error_word) {}
  __device__ __forceinline__ bool expired() {
    if (++it < 256) return false;
    if ((it & 63u) != 0) return false;
    if (err && (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 8u)) return true;
    const uint64_t now = wall_clock64();
    if (t0 == 0) {
      t0 = now;
      return false;
    }
    return now - t0 > kSpinTicks;
  }
};

// ------------------------------------------------------------------ wave primitives
__device__ __forceinline__ uint32_t lane() { return __lane_id(); }
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
// Exclusive prefix sum over the 64 lanes of a fully active wave; *total gets the wave's sum. An
// inclusive scan with DPP moves (row shifts by 1, 2, 4, 8 inside each 16-lane row, then the row
// broadcasts of lanes 15 and 31 into the rows above), which stay in the VALU instead of going
// through the LDS crossbar as lane shuffles do.
__device__ __forceinline__ uint32_t wave_exclusive_sum(uint32_t v, uint32_t* total) {
  uint32_t x = v;
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31
  *total = __builtin_amdgcn_readlane(x, 63);
  return x - v;
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ uint64_t active() { return __ballot(1); }
__device__ __forceinline__ uint64_t lt_mask() { return (1ull << lane()) - 1ull; }
__device__ __forceinline__ int first(uint64_t m) { return __ffsll((unsigned long long)m) - 1; }
__device__ __forceinline__ uint32_t popc(uint64_t m) { return (uint32_t)__popcll(m); }
__device__ __forceinline__ uint32_t rank_in(uint64_t m) { return popc(m & lt_mask()); }

__device__ __forceinline__ uint32_t readlane(uint32_t v, int l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l) {
  uint32_t lo = readlane((uint32_t)v, l), hi = readlane((uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}
// Per-lane source broadcast (ds_bpermute). src must be an active lane.
__device__ __forceinline__ uint32_t shfl(uint32_t v, int src) {
  return (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)v);
}
__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
  uint32_t lo = shfl((uint32_t)v, src), hi = shfl((uint32_t)(v >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}

// Mask of the W-lane segment that contains lane l (W = 64 covers the whole wave).
template <int W>
__device__ __forceinline__ uint64_t segment_of(int l) {
  if constexpr (W >= 64) return ~0ull;
  else return ((1ull << W) - 1ull) << ((l / W) * W);
}

// 64-lane match_any: returns the mask of active lanes inside this lane's W-lane segment whose
// key equals this lane's key. Replaces __match_any_sync (Utility.cuh:80, BulkSemaphore_impl.cuh:124).
// Iterates once per distinct (segment, key) pair; `todo` stays wave-uniform.
template <int W>
__device__ __forceinline__ uint64_t match(uint64_t key) {
  uint64_t todo = active();
  uint64_t mine = 0;
  while (todo) {
    const int l = first(todo);
    const uint64_t k = readlane64(key, l);
    const uint64_t grp = ballot(key == k) & todo & segment_of<W>(l);
    if ((grp >> lane()) & 1ull) mine = grp;
    todo &= ~grp;
  }
  return mine;
}

// Position of the n-th (0-based) set bit of m; m must have more than n bits set.
__device__ __forceinline__ uint32_t nth_set_bit(uint64_t m, uint32_t n) {
  uint32_t pos = 0;
#pragma unroll
  for (int width = 32; width >= 1; width >>= 1) {
    const uint64_t low = m & ((width == 64) ? ~0ull : ((1ull << width) - 1ull));
    const uint32_t c = popc(low);
    if (n >= c) {
      n -= c;
      m >>= width;
      pos += width;
    } else {
      m = low;
    }
  }
  return pos;
}

// Lowest `n` set bits of m (n may exceed popc(m)).
__device__ __forceinline__ uint64_t lowest_bits(uint64_t m, uint32_t n) {
  if (n >= popc(m)) return m;
  if (n == 0) return 0;
  const uint32_t top = nth_set_bit(m, n - 1);
  return (top == 63) ? m : (m & ((2ull << top) - 1ull));
}

}  // namespace dev
}  // namespace ouro
